/*
 * Marching-cubes case table and cube geometry shared by the HIP kernels and
 * the CPU oracle.
 *
 * The triangle table is the classic public-domain Lorensen/Bourke table, the
 * same one the reference ships as MarchingCubes::triTable
 * (src/mc_tables.h:17-273).  Triangle order and winding must be identical for
 * index-buffer parity, so the data is the same; the storage is ours: one
 * 64-bit word per case, the triangle count in bits 60..63 and the edge of
 * triangle k, vertex j in nibble 3k+j.  A wavefront reads one 8-byte word per
 * cell instead of sixteen int32 entries.
 *
 * The edge table (reference src/mc_tables.h:275-308) is not stored: edge e of
 * case c is active iff the two corners of e are on different sides of the
 * isovalue, which is what mtx_case_edge_mask computes.
 *
 * Corner numbering follows the reference shaders
 * (src/shaders/marching_cubes_pmb.task.glsl:81-84): corner i sits at
 * offset (x,y,z) with x = ((i+1)>>1)&1, y = (i>>1)&1, z = i>>2.
 */
#ifndef MTX_MC_CASES_H
#define MTX_MC_CASES_H

#include <stdint.h>

#if defined(__HIPCC__)
#define MTX_HD __host__ __device__
#else
#define MTX_HD
#endif
/* constexpr where the language has it (the oracle includes this from C) */
#ifdef __cplusplus
#define MTX_CE constexpr
#else
#define MTX_CE
#endif

static const uint64_t mtx_tri_words[256] = {
    0x0000000000000000ull, 0x1000000000000380ull, 0x1000000000000910ull, 0x2000000000189381ull,
    0x1000000000000a21ull, 0x2000000000a21380ull, 0x2000000000920a29ull, 0x300000089a8a2382ull,
    0x10000000000002b3ull, 0x20000000000b82b0ull, 0x2000000000b32091ull, 0x3000000b89b912b1ull,
    0x20000000003ab1a3ull, 0x3000000ab8a801a0ull, 0x30000009ab9b3093ull, 0x2000000000b8aa89ull,
    0x1000000000000874ull, 0x2000000000437034ull, 0x2000000000748910ull, 0x3000000137174914ull,
    0x2000000000748a21ull, 0x3000000a21403743ull, 0x3000000748209a29ull, 0x40004973727929a2ull,
    0x20000000002b3748ull, 0x300000040242b74bull, 0x3000000b32748109ull, 0x40001292b9b49b74ull,
    0x3000000487ab31a3ull, 0x40004b7401b41ab1ull, 0x400030bab9b09874ull, 0x3000000ab99b4b74ull,
    0x1000000000000459ull, 0x2000000000380459ull, 0x2000000000051450ull, 0x3000000513538458ull,
    0x2000000000459a21ull, 0x3000000594a21803ull, 0x3000000204245a25ull, 0x40008434535235a2ull,
    0x2000000000b32459ull, 0x3000000594b802b0ull, 0x3000000b32510450ull, 0x4000584b82852512ull,
    0x300000045931ab3aull, 0x4000ab81a8180594ull, 0x400030bab5b05045ull, 0x3000000b8aa85845ull,
    0x2000000000975879ull, 0x3000000375359039ull, 0x3000000751710870ull, 0x2000000000753351ull,
    0x300000021a759879ull, 0x400037503505921aull, 0x400025a758528208ull, 0x30000007533525a2ull,
    0x30000002b3987597ull, 0x4000b72029279759ull, 0x4000751871810b32ull, 0x300000051771b12bull,
    0x4000b3a31a758859ull, 0x50aba010b7905075ull, 0x507570805a30b0abull, 0x20000000005b75abull,
    0x100000000000056aull, 0x20000000006a5380ull, 0x20000000006a5109ull, 0x30000006a5891381ull,
    0x2000000000162561ull, 0x3000000803621561ull, 0x3000000620609569ull, 0x4000823625285895ull,
    0x200000000056ab32ull, 0x300000056a02b80bull, 0x30000006a5b32910ull, 0x4000b892b92916a5ull,
    0x3000000315356b36ull, 0x40006b51505b0b80ull, 0x40009505606306b3ull, 0x300000089bb96956ull,
    0x20000000008746a5ull, 0x3000000a56374034ull, 0x30000007486a5091ull, 0x400049737179156aull,
    0x3000000874156216ull, 0x4000743403625521ull, 0x4000620560509748ull, 0x5962695923497937ull,
    0x300000056a4872b3ull, 0x4000b720242746a5ull, 0x40006a5b32874910ull, 0x56a54b7b492b9129ull,
    0x40006b51535b3748ull, 0x5b404b7b016b5b15ull, 0x574836b630560950ull, 0x40009b7974b96956ull,
    0x2000000000a4694aull, 0x3000000380a946a4ull, 0x300000004606a10aull, 0x4000a16468618138ull,
    0x3000000462421941ull, 0x4000462942921803ull, 0x2000000000624420ull, 0x3000000624428238ull,
    0x300000032b46a94aull, 0x40006a4a94b82280ull, 0x4000a164606102b3ull, 0x51b8b12184a16146ull,
    0x400036b319639469ull, 0x514641916b0181b8ull, 0x30000004600636b3ull, 0x200000000086b846ull,
    0x3000000a98a876a7ull, 0x4000a76a907a0370ull, 0x40000818717a176aull, 0x300000037117a76aull,
    0x4000768981861621ull, 0x5937390976192962ull, 0x3000000206607087ull, 0x2000000000276237ull,
    0x400076898a86ab32ull, 0x57a9a76790b72702ull, 0x5b32a767a1871081ull, 0x400017616a71b12bull,
    0x563136b619768698ull, 0x200000000076b190ull, 0x400006b0b3607087ull, 0x10000000000006b7ull,
    0x1000000000000b67ull, 0x200000000067b803ull, 0x200000000067b910ull, 0x300000067b138918ull,
    0x20000000007b621aull, 0x30000007b6803a21ull, 0x30000007b69a2092ull, 0x400089a38a3a27b6ull,
    0x2000000000726327ull, 0x3000000026067807ull, 0x3000000910732672ull, 0x4000678891681261ull,
    0x300000073171a67aull, 0x4000801781a7167aull, 0x40007a69a0a70730ull, 0x30000009a88a7a67ull,
    0x200000000068b486ull, 0x3000000640603b63ull, 0x3000000109648b68ull, 0x400063b139369649ull,
    0x30000001a28b6486ull, 0x4000640b60b03a21ull, 0x40009a2920b648b4ull, 0x536463b34923a39aull,
    0x3000000264248328ull, 0x2000000000264240ull, 0x4000834642432091ull, 0x3000000642241491ull,
    0x40001a6648168318ull, 0x300000040660a01aull, 0x539a9303a6834364ull, 0x20000000004a649aull,
    0x2000000000b67594ull, 0x300000067b594380ull, 0x3000000b67045105ull, 0x400051345343867bull,
    0x3000000b6721a459ull, 0x4000594380a217b6ull, 0x4000204a24a45b67ull, 0x567b25a523453843ull,
    0x3000000945267327ull, 0x4000786260680459ull, 0x4000045051673263ull, 0x5851584812786826ull,
    0x400073167161a459ull, 0x5459078701671a61ull, 0x5a737a6a305a4a04ull, 0x4000a84a458a7a67ull,
    0x300000098b9b6596ull, 0x4000590650360b63ull, 0x4000b65510b508b0ull, 0x30000001355363b6ull,
    0x400065b8b9b59a21ull, 0x5a21965690b603b0ull, 0x552025a50865b58bull, 0x400035a3a25363b6ull,
    0x4000283265825985ull, 0x3000000260069659ull, 0x5826283865081851ull, 0x2000000000612651ull,
    0x5698965683a61631ull, 0x400006505960a01aull, 0x2000000000a65830ull, 0x100000000000065aull,
    0x2000000000b57a5bull, 0x300000003857ba5bull, 0x3000000091ba57b5ull, 0x40001381897ba57aull,
    0x300000015717b21bull, 0x4000b27571721380ull, 0x40007b2209729579ull, 0x5289823295b27257ull,
    0x3000000573532a52ull, 0x400052a578258028ull, 0x40002a37353a5109ull, 0x525752a278129289ull,
    0x2000000000573531ull, 0x3000000571170780ull, 0x3000000735539309ull, 0x2000000000795789ull,
    0x30000008ba8a5485ull, 0x400003bba50b5405ull, 0x400054aba8a48910ull, 0x541314943b54a4baull,
    0x40008548b2582152ull, 0x5b151b2b543b0b40ull, 0x558b8545b2950520ull, 0x20000000003b2549ull,
    0x4000483543253a52ull, 0x30000000244252a5ull, 0x5910854583a532a3ull, 0x40002492914252a5ull,
    0x3000000153358548ull, 0x2000000000501540ull, 0x4000530509358548ull, 0x1000000000000549ull,
    0x3000000ba9b947b4ull, 0x4000ba97b9794380ull, 0x4000b470414b1ba1ull, 0x54bab474a1843413ull,
    0x4000219b294b97b4ull, 0x53801b2b197b9479ull, 0x300000004224b47bull, 0x400042343824b47bull,
    0x4000947732972a92ull, 0x570207872a4797a9ull, 0x5a040a1a472a3a73ull, 0x20000000004782a1ull,
    0x3000000317714194ull, 0x4000178180714194ull, 0x2000000000347304ull, 0x1000000000000784ull,
    0x20000000008ba8a9ull, 0x3000000a9bb93903ull, 0x3000000ba88a0a10ull, 0x2000000000a3ba13ull,
    0x30000008b99b1b21ull, 0x40009b2921b93903ull, 0x2000000000b08b20ull, 0x1000000000000b23ull,
    0x300000098aa82832ull, 0x20000000002902a9ull, 0x40008a1810a82832ull, 0x10000000000002a1ull,
    0x2000000000819831ull, 0x1000000000000190ull, 0x1000000000000830ull, 0x0000000000000000ull,
};

/* Number of triangles case c produces (0..5). */
static inline MTX_CE MTX_HD unsigned mtx_word_ntri(uint64_t w) { return (unsigned)(w >> 60); }
/* Edge index (0..11) of vertex j of triangle k. */
static inline MTX_CE MTX_HD unsigned mtx_word_edge(uint64_t w, unsigned k, unsigned j)
{
    return (unsigned)((w >> (4u * (3u * k + j))) & 0xFu);
}

/* Corner offsets, bit-packed: corner i -> (x,y,z). */
static inline MTX_CE MTX_HD unsigned mtx_corner_x(unsigned i) { return ((i + 1u) >> 1) & 1u; }
static inline MTX_CE MTX_HD unsigned mtx_corner_y(unsigned i) { return (i >> 1) & 1u; }
static inline MTX_CE MTX_HD unsigned mtx_corner_z(unsigned i) { return i >> 2; }

/* Cube edges: edges 0..3 run around the z=0 face (0-1, 1-2, 2-3, 3-0),
 * edges 4..7 around the z=1 face (4-5, 5-6, 6-7, 7-4), edges 8..11 are the
 * vertical edges i -> i+4. */
static inline MTX_CE MTX_HD unsigned mtx_edge_c0(unsigned e) { return e < 8u ? e : e - 8u; }
static inline MTX_CE MTX_HD unsigned mtx_edge_c1(unsigned e)
{
    return e < 8u ? ((e & 4u) | ((e + 1u) & 3u)) : e - 4u;
}

/* Owner of an edge under PMB ownership: the cell whose minimum corner is the
 * edge's lower endpoint, and the axis (0=x, 1=y, 2=z) the edge runs along.
 * Packed as dx | dy<<1 | dz<<2 | axis<<3.  Derived from the corner geometry;
 * equals the reference's edgeOwner table
 * (src/shaders/marching_cubes_pmb.mesh.glsl:39-42). */
static inline MTX_CE MTX_HD unsigned mtx_edge_owner(unsigned e)
{
    unsigned a = mtx_edge_c0(e), b = mtx_edge_c1(e);
    unsigned ax = mtx_corner_x(a), ay = mtx_corner_y(a), az = mtx_corner_z(a);
    unsigned bx = mtx_corner_x(b), by = mtx_corner_y(b), bz = mtx_corner_z(b);
    unsigned dx = ax < bx ? ax : bx, dy = ay < by ? ay : by, dz = az < bz ? az : bz;
    unsigned axis = (ax != bx) ? 0u : ((ay != by) ? 1u : 2u);
    return dx | (dy << 1) | (dz << 2) | (axis << 3);
}

/* Active-edge mask of case c (bit e set iff edge e crosses the isovalue). */
static inline MTX_HD unsigned mtx_case_edge_mask(unsigned c)
{
    unsigned m = 0;
    for (unsigned e = 0; e < 12u; ++e)
        m |= (((c >> mtx_edge_c0(e)) ^ (c >> mtx_edge_c1(e))) & 1u) << e;
    return m;
}

/* Owner index of an edge, 0..11: u = 4 axis + d_p + 2 d_q, where (p, q) are
 * the two axes other than the edge's, in increasing order, and d its owner
 * corner's offset from the cell's minimum corner.  A cell's meshlet vertices
 * are taken in ascending u (the order of owner corners along the axes). */
static inline MTX_CE MTX_HD unsigned mtx_edge_uidx(unsigned e)
{
    unsigned o = mtx_edge_owner(e);
    unsigned a = o >> 3;
    unsigned p = a == 0u ? 1u : 0u, q = a == 2u ? 1u : 2u;
    return 4u * a + ((o >> p) & 1u) + 2u * ((o >> q) & 1u);
}

/* The edge of owner index u (inverse of mtx_edge_uidx). */
static inline MTX_CE MTX_HD unsigned mtx_uidx_edge(unsigned u)
{
    unsigned e = 0;
    for (unsigned k = 0; k < 12u; ++k)
        if (mtx_edge_uidx(k) == u) e = k;
    return e;
}

/* Active edges of case c as owner-index bits (bit u). */
static inline MTX_HD unsigned mtx_case_umask(unsigned c)
{
    unsigned m = 0;
    for (unsigned e = 0; e < 12u; ++e)
        m |= (((c >> mtx_edge_c0(e)) ^ (c >> mtx_edge_c1(e))) & 1u) << mtx_edge_uidx(e);
    return m;
}

#endif /* MTX_MC_CASES_H */
