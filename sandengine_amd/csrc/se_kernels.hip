// Kernel templates for one sandengine grid step on CDNA4 (gfx950).
//
// This text is appended to the per-rule-set code the HIP backend generates
// (se_hipgen.cpp: SE_NMAT, se_fixed(), se_block(...)) and JIT-compiled with
// hiprtc.  It is not compiled on its own.
//
// Grid layout in HBM: u32 cell ids, row-major, row pitch `pitch` cells
// (a multiple of 64, so every row starts 256-byte aligned); x >= W is padding
// that is never read as a cell.  Out-of-bounds cells read as WALL
// (shaders/compute/operations.glsl:45-51 with SCREEN_IS_BORDER); ids >= SE_NMAT
// decode to NULL (gen/materials.glsl:79-86).
//
// The step: Margolus 2x2 blocks with origin 2*i - off, off = offset(frame % 4)
// (operations.glsl:25-34); one evaluation per block is exactly equivalent to
// the reference's one invocation per cell (falling_sand.glsl:61-118, each
// invocation recomputes its block and keeps one cell).
typedef unsigned int u32;
typedef unsigned long long u64;

#define SE_EMPTY 0u
#define SE_NULLID 1u
#define SE_WALLID 2u

struct se_mod {  // == SimModification (sandengine-core/src/simulation.rs:46-56), 32 B
    int x, y, shape, size, mat, pad0, pad1, pad2;
};

__device__ __forceinline__ u32 se_decode(u32 v) { return v < SE_NMAT ? v : SE_NULLID; }

// sqrt_rn(s) <= R for f32 s >= 0, exactly: the round-to-nearest square root is
// <= R iff sqrt(s) < R + ulp(R)/2 (no ties are possible for float s), and that
// midpoint squared is exact in double.
__device__ __forceinline__ bool se_sqrt_le(float s, float R) {
    if (!(R >= 0.0f)) return false;
    if (R == __builtin_inff()) return true;
    const float up = __builtin_bit_cast(float, __builtin_bit_cast(unsigned int, R) + 1u);  // next float above R
    double m = 0.5 * ((double)R + (double)up);
    return (double)s < m * m;
}

// Modification override for cell (x, y): returns the id or -1
// (falling_sand.glsl:134-179: last matching entry wins, a NULL material hides
// earlier matches, the list was cut at the first size-0 entry by the host).
__device__ __forceinline__ int se_mod_override(const se_mod* __restrict__ mods, int nmods, int x, int y) {
    int fin = -1;
    for (int i = 0; i < nmods; i++) {
        const int mx = mods[i].x, my = mods[i].y, sz = mods[i].size, sh = mods[i].shape;
        int dx = mx - x, dy = my - y;
        dx = dx < 0 ? -dx : dx;
        dy = dy < 0 ? -dy : dy;
        bool hit;
        if (sh == 0) {
            float fx = (float)dx, fy = (float)dy;
            float s = __fadd_rn(__fmul_rn(fx, fx), __fmul_rn(fy, fy));
            hit = se_sqrt_le(s, (float)sz);
        } else if (sh == 1) {
            hit = dx <= sz && dy <= sz;
        } else {
            hit = false;
        }
        if (hit) {
            const int m = mods[i].mat;
            fin = (m >= 0 && (u32)m < SE_NMAT) ? m : (int)SE_NULLID;
        }
    }
    return fin == (int)SE_NULLID ? -1 : fin;
}

// Does any modification's bounding box touch rows [ya, yb] x columns [xa, xb]?
__device__ __forceinline__ bool se_mods_touch(const se_mod* __restrict__ mods, int nmods, int xa, int xb, int ya,
                                              int yb) {
    for (int i = 0; i < nmods; i++) {
        const long r = mods[i].size < 0 ? 0 : mods[i].size;
        const long mx = mods[i].x, my = mods[i].y;
        if (mx + r >= xa && mx - r <= xb && my + r >= ya && my - r <= yb) return true;
    }
    return false;
}

__device__ __forceinline__ u32 se_cell_at(const u32* __restrict__ row, int x, int W) {
    return x < W ? se_decode(row[x]) : SE_WALLID;
}

// ---------------------------------------------------------------------------
// Plain kernel: one wavefront per (256-cell strip, block row); each lane owns an
// aligned quad of 4 cells in both rows of the block row and loads/stores it with
// 16-byte accesses.  With x offset 0 the lane's two blocks lie inside its quad.
// With x offset 1 the lane evaluates the blocks at x0+1 and x0+3; the cell x0+4
// comes from lane+1 and the new value of x0 from lane-1 (cross-lane shuffles),
// lane 63 loads/stores the one cell past the strip, and lane 0 of strip 0
// evaluates the edge block at x = -1.  Every block is evaluated exactly once.
// Reads `in`, writes `out` (ping-pong).
// ---------------------------------------------------------------------------
typedef unsigned int se_u32x4 __attribute__((ext_vector_type(4)));

// 16-byte non-temporal load: every input line is read once per step (measured
// +1.5 % on the plain kernel over ordinary loads)
__device__ __forceinline__ uint4 se_load_nt(const u32* p) {
    const se_u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const se_u32x4*>(p));
    return make_uint4(v.x, v.y, v.z, v.w);
}

__device__ __forceinline__ uint4 se_load_quad(const u32* __restrict__ row, int x0, int W) {
    uint4 q = make_uint4(SE_WALLID, SE_WALLID, SE_WALLID, SE_WALLID);
    if (row != nullptr && x0 < W) {
        uint4 v = se_load_nt(row + x0);
        q.x = se_decode(v.x);
        q.y = x0 + 1 < W ? se_decode(v.y) : SE_WALLID;
        q.z = x0 + 2 < W ? se_decode(v.z) : SE_WALLID;
        q.w = x0 + 3 < W ? se_decode(v.w) : SE_WALLID;
    }
    return q;
}

__device__ __forceinline__ void se_apply_mods_quad(uint4& q, const se_mod* __restrict__ mods, int nmods, int x0,
                                                   int y) {
    int o;
    if ((o = se_mod_override(mods, nmods, x0, y)) >= 0) q.x = (u32)o;
    if ((o = se_mod_override(mods, nmods, x0 + 1, y)) >= 0) q.y = (u32)o;
    if ((o = se_mod_override(mods, nmods, x0 + 2, y)) >= 0) q.z = (u32)o;
    if ((o = se_mod_override(mods, nmods, x0 + 3, y)) >= 0) q.w = (u32)o;
}


template <bool NT>
__device__ __forceinline__ void se_store_quad(u32* p, const uint4& q) {
    if (NT) {
        se_u32x4 v = {q.x, q.y, q.z, q.w};
        __builtin_nontemporal_store(v, reinterpret_cast<se_u32x4*>(p));
    } else {
        *reinterpret_cast<uint4*>(p) = q;
    }
}

typedef float se_f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void se_store_f4(float4* p, const float4& v) {
    se_f32x4 t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<se_f32x4*>(p));
}

// One wavefront handles R consecutive block rows of one strip: all 2R row loads
// are issued before any block is evaluated, to keep more bytes in flight.
template <int R, bool NT>
__device__ __forceinline__ void se_plain_body(const u32* __restrict__ in, u32* __restrict__ out, int W, int H,
                                              int pitch, int frame, int offx, int offy, int strips, int brows,
                                              const se_mod* __restrict__ mods, int nmods) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    const int strip = wave % strips;
    const int brow0 = (wave / strips) * R;
    if (brow0 >= brows) return;  // whole wavefront leaves together
    const int base = strip * 256;
    const int x0 = base + 4 * lane;
    const u32 fr = (u32)frame;
    uint4 a[R], b[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ya = 2 * (brow0 + r) - offy, yb = ya + 1;
        const bool live = brow0 + r < brows;
        a[r] = se_load_quad((live && ya >= 0 && ya < H) ? in + (size_t)ya * pitch : nullptr, x0, W);
        b[r] = se_load_quad((live && yb < H) ? in + (size_t)yb * pitch : nullptr, x0, W);
    }
    u32 ea[R], eb[R];
    if (offx == 1) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            // cell x0+4 of both rows: lane+1's first cell; lane 63 loads it
            ea[r] = __shfl_down(a[r].x, 1, 64);
            eb[r] = __shfl_down(b[r].x, 1, 64);
            if (lane == 63) {
                const int ya = 2 * (brow0 + r) - offy, yb = ya + 1;
                const bool live = brow0 + r < brows;
                const int xe = base + 256;
                ea[r] = (live && ya >= 0 && ya < H) ? se_cell_at(in + (size_t)ya * pitch, xe, W) : SE_WALLID;
                eb[r] = (live && yb < H) ? se_cell_at(in + (size_t)yb * pitch, xe, W) : SE_WALLID;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int ya = 2 * (brow0 + r) - offy, yb = ya + 1;
        const bool live = brow0 + r < brows;
        u32 p0 = a[r].x, p1 = a[r].y, p2 = a[r].z, p3 = a[r].w;
        u32 q0 = b[r].x, q1 = b[r].y, q2 = b[r].z, q3 = b[r].w;
        if (offx == 0) {
            se_block(p0, p1, q0, q1, (u32)x0, (u32)ya, fr);
            se_block(p2, p3, q2, q3, (u32)(x0 + 2), (u32)ya, fr);
        } else {
            u32 e0 = ea[r], e1 = eb[r];
            u32 edge_a = 0, edge_b = 0;
            if (strip == 0 && lane == 0) {  // block at x = -1: left column is WALL
                u32 w0 = SE_WALLID, w2 = SE_WALLID;
                edge_a = p0;
                edge_b = q0;
                se_block(w0, edge_a, w2, edge_b, 0xFFFFFFFFu, (u32)ya, fr);
            }
            se_block(p1, p2, q1, q2, (u32)(x0 + 1), (u32)ya, fr);
            se_block(p3, e0, q3, e1, (u32)(x0 + 3), (u32)ya, fr);
            // new value of cell x0 comes from lane-1's second block
            u32 na = __shfl_up(e0, 1, 64), nb = __shfl_up(e1, 1, 64);
            if (lane == 0) {
                na = edge_a;
                nb = edge_b;
            }
            p0 = na;
            q0 = nb;
            if (lane == 63 && live && base + 256 < W) {  // the cell past the strip
                const int xe = base + 256;
                if (nmods > 0) {
                    int o;
                    if ((o = se_mod_override(mods, nmods, xe, ya)) >= 0) e0 = (u32)o;
                    if ((o = se_mod_override(mods, nmods, xe, yb)) >= 0) e1 = (u32)o;
                }
                if (ya >= 0) out[(size_t)ya * pitch + xe] = e0;
                if (yb < H) out[(size_t)yb * pitch + xe] = e1;
            }
        }
        if (!live || x0 >= W) continue;
        uint4 na = make_uint4(p0, p1, p2, p3), nb = make_uint4(q0, q1, q2, q3);
        if (nmods > 0 && se_mods_touch(mods, nmods, base, base + 256, ya, yb)) {
            se_apply_mods_quad(na, mods, nmods, x0, ya);
            se_apply_mods_quad(nb, mods, nmods, x0, yb);
        }
        if (offx == 1 && lane == 0 && strip != 0) {
            // cell x0 belongs to the previous strip's last block: store x0+1..x0+3
            if (ya >= 0) {
                u32* p = out + (size_t)ya * pitch + x0;
                p[1] = na.y; p[2] = na.z; p[3] = na.w;
            }
            if (yb < H) {
                u32* p = out + (size_t)yb * pitch + x0;
                p[1] = nb.y; p[2] = nb.z; p[3] = nb.w;
            }
            continue;
        }
        if (ya >= 0) se_store_quad<NT>(out + (size_t)ya * pitch + x0, na);
        if (yb < H) se_store_quad<NT>(out + (size_t)yb * pitch + x0, nb);
    }
}

#define SE_PLAIN_KERNEL(R, NT)                                                                                   \
    extern "C" __global__ __launch_bounds__(256) void se_step_plain_r##R##_nt##NT(                              \
        const u32* __restrict__ in, u32* __restrict__ out, int W, int H, int pitch, int frame, int offx,       \
        int offy, int strips, int brows, const se_mod* __restrict__ mods, int nmods) {                         \
        se_plain_body<R, NT != 0>(in, out, W, H, pitch, frame, offx, offy, strips, brows, mods, nmods);         \
    }
SE_PLAIN_KERNEL(1, 0)
SE_PLAIN_KERNEL(2, 0)
SE_PLAIN_KERNEL(4, 0)
SE_PLAIN_KERNEL(1, 1)
SE_PLAIN_KERNEL(2, 1)
SE_PLAIN_KERNEL(4, 1)

// Streaming copy of n4 16-byte quads (the HBM reference for the plain kernel:
// the same bytes read and written, no rules).
extern "C" __global__ __launch_bounds__(256) void se_copy(const u32* __restrict__ src, u32* __restrict__ dst,
                                                          size_t n4) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256)
        se_store_quad<true>(dst + 4 * i, se_load_nt(src + 4 * i));
}

// Frame 1 of a world clears every cell to EMPTY (falling_sand.glsl:128-131).
extern "C" __global__ __launch_bounds__(256) void se_fill(u32* __restrict__ g, size_t n, u32 v) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) g[i] = v;
}

// ---------------------------------------------------------------------------
// Cell codes in LDS.  With the coded table (SE_TBL == 2) the LDS tile holds
// code(id) = (id - 3) mod SE_NMAT: the rule set's materials 3.. become 0..,
// EMPTY becomes D - 1, NULL D and WALL D + 1 (D = SE_NMAT - 2), so the tuples the
// table covers are those of codes < D.  With the full table (SE_TBL == 1) the
// LDS tile holds id * SE_CS, SE_CS = 4R (R hash classes) where that is a power
// of two and keeps every code in a byte: a block's tuple taken over such codes
// is then the byte offset of its first table entry (one instruction fewer per
// block than scaling the tuple).  The chain path keeps ids in LDS.
// se_enc_ok takes an id < SE_NMAT; se_enc decodes any u32 first.
// ---------------------------------------------------------------------------
#if SE_TBL == 1
#define SE_CS_4R (SE_TBL_R * 4u)
#define SE_CS (((SE_CS_4R & (SE_CS_4R - 1u)) == 0u && SE_CS_4R * (SE_NMAT - 1u) <= 255u) ? SE_CS_4R : 1u)
#else
#define SE_CS 1u
#endif
// log2 of SE_CS (a power of two)
#define SE_CS_SHIFT (SE_CS == 1u ? 0u : SE_CS == 2u ? 1u : SE_CS == 4u ? 2u : SE_CS == 8u ? 3u : SE_CS == 16u ? 4u \
                     : SE_CS == 32u ? 5u : SE_CS == 64u ? 6u : 7u)
#if SE_TBL == 2
#define SE_LWALL (SE_NMAT - 1u)
__device__ __forceinline__ u32 se_enc_ok(u32 id) { return min(id - 3u, id + (SE_NMAT - 3u)); }
__device__ __forceinline__ u32 se_dec(u32 c) { return min(c + 3u, c - (SE_NMAT - 3u)); }
#else
#define SE_LWALL (SE_WALLID * SE_CS)
__device__ __forceinline__ u32 se_enc_ok(u32 id) { return id << SE_CS_SHIFT; }
__device__ __forceinline__ u32 se_dec(u32 c) { return c >> SE_CS_SHIFT; }
#endif
__device__ __forceinline__ u32 se_enc(u32 v) { return se_enc_ok(se_decode(v)); }

// ---------------------------------------------------------------------------
// Transition table (SE_TBL == 1): entry tuple * SE_TBL_R + r holds the block
// result (four ids packed as bytes c0 | c1<<8 | c2<<16 | c3<<24) for ids
// tuple = c0 + N*(c1 + N*(c2 + N*c3)) and hash-threshold class r.  Built on the
// device from se_block_h itself with one representative hash per class.
// ---------------------------------------------------------------------------
// 1 when no entry of the full table turns a WALL cell into something else (set
// by the host after checking the built table): edge tiles then skip the
// restoring of WALL outside the world, which the table already keeps.
extern "C" __device__ u32 se_wall_kept = 0u;

// Bit c set (by the host, from the built table) when a block of four cells of
// id c (full table) or LDS code c (coded table) comes out unchanged under every
// hash class and mirror bit.  A tile whose cells all hold such a c is a fixed
// point of every sub-step: each block evaluated in it holds four cells of c
// (se_lds_body skips its sub-steps).
// 0 (never skip) for rule sets without a table.
extern "C" __device__ u32 se_stable_mask = 0u;

extern "C" __global__ __launch_bounds__(256) void se_build_table(u32* __restrict__ tbl) {
    const u32 tuple = blockIdx.x * 256u + threadIdx.x;
    if (tuple >= SE_TBL_ENTRIES / SE_TBL_R) return;
    u32 t = tuple;
    const u32 i0 = t % SE_NMAT;
    t /= SE_NMAT;
    const u32 i1 = t % SE_NMAT;
    t /= SE_NMAT;
    const u32 i2 = t % SE_NMAT;
    const u32 i3 = t / SE_NMAT;
    for (u32 r = 0; r < SE_TBL_R; r++) {
        u32 h[4], rr = r;
#pragma unroll
        for (int l = 0; l < 4; l++) {
            const u32 cnt = se_rclass_count(l);
            h[l] = se_rclass_rep(l, rr % cnt);
            rr /= cnt;
        }
        u32 c0 = i0, c1 = i1, c2 = i2, c3 = i3;
        se_block_h(c0, c1, c2, c3, h[0], h[1], h[2], h[3]);
        tbl[tuple * SE_TBL_R + r] = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
    }
}

#if SE_TBL == 2
// Coded table, raw form for the host, which names the distinct outcomes and
// packs the LDS table: for each tuple of codes < D (index c0 + D*(c1 + D*(c2 +
// D*c3))), each class r of hash lanes 1..3 and each mirror bit m, the block
// result as ids packed in bytes, from se_block_h itself with one representative
// hash per class (lane 0 at the mirror threshold for m = 1, above it for m = 0).
extern "C" __global__ __launch_bounds__(256) void se_build_ctable(u32* __restrict__ raw) {
    const u32 tuple = blockIdx.x * 256u + threadIdx.x;
    if (tuple >= SE_CT_D * SE_CT_D * SE_CT_D * SE_CT_D) return;
    u32 t = tuple;
    const u32 i0 = se_dec(t % SE_CT_D);
    t /= SE_CT_D;
    const u32 i1 = se_dec(t % SE_CT_D);
    t /= SE_CT_D;
    const u32 i2 = se_dec(t % SE_CT_D);
    const u32 i3 = se_dec(t / SE_CT_D);
    for (u32 r = 0; r < SE_CT_R; r++) {
        u32 h[4], rr = r;
#pragma unroll
        for (int l = 1; l < 4; l++) {
            const u32 cnt = se_rclass_count(l);
            h[l] = se_rclass_rep(l, rr % cnt);
            rr /= cnt;
        }
        for (u32 m = 0; m < 2; m++) {
            h[0] = m ? SE_CT_MIRROR_T : 0xffffffffu;
            u32 c0 = i0, c1 = i1, c2 = i2, c3 = i3;
            se_block_h(c0, c1, c2, c3, h[0], h[1], h[2], h[3]);
            raw[(tuple * SE_CT_R + r) * 2u + m] = c0 | (c1 << 8) | (c2 << 16) | (c3 << 24);
        }
    }
}
#endif

// ---------------------------------------------------------------------------
// LDS temporal-blocked kernel: a workgroup of NT threads stages a
// (TW + 2*HB) x (TH + 2*HB) tile (u8 ids, or u8 codes with the coded table) in
// LDS, stages the rule set's table (if any) after it, and advances it `nsub` <= HB
// Margolus sub-steps in place (blocks partition the cells, so in-place is
// race-free); each sub-step evaluates only blocks whose four cells are still
// valid, so the valid region shrinks by at most one cell per side per sub-step
// and the TW x TH interior is exact after HB sub-steps.  Cells outside the world
// are WALL in LDS and are never written.  Modifications apply to sub-step 0
// only.  Reads `in`, writes the interior to `out` (ping-pong).  TW, TH, HB are
// multiples of 4, so tile rows start 4-byte aligned in LDS and 16-byte aligned
// in HBM, and the tile moves as 4-cell quads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void se_offset(u32 f, int& ox, int& oy) {
    const u32 m = f & 3u;  // frame % 4 for frame >= 0
    ox = (m == 1u || m == 3u) ? 1 : 0;
    oy = (m == 1u || m == 2u) ? 1 : 0;
}

__device__ __forceinline__ u32 se_pack4(u32 a, u32 b, u32 c, u32 d) { return a | (b << 8) | (c << 16) | (d << 24); }

// Four cells from HBM as LDS bytes, ids >= SE_NMAT decoded to NULL (one range
// test for the quad; the per-cell decode runs only where one is out of range).
// Byte-wise (b + a) mod SE_NMAT of packed bytes b < SE_NMAT, a < SE_NMAT (the
// coded table has SE_NMAT <= 18, so no byte carries into the next).
__device__ __forceinline__ u32 se_rot4(u32 p, u32 a) {
    const u32 q = p + a * 0x01010101u;
    const u32 ge = (q + (0x80u - SE_NMAT) * 0x01010101u) & 0x80808080u;  // bytes q >= SE_NMAT
    return q - (ge >> 7) * SE_NMAT;
}

__device__ __forceinline__ u32 se_pack_quad(uint4 v) {
    const u32 mx = max(max(v.x, v.y), max(v.z, v.w));
    if (__builtin_expect(mx >= SE_NMAT, 0)) return se_pack4(se_enc(v.x), se_enc(v.y), se_enc(v.z), se_enc(v.w));
#if SE_TBL == 2
    return se_rot4(se_pack4(v.x, v.y, v.z, v.w), SE_NMAT - 3u);  // codes (id - 3) mod SE_NMAT
#else
    return se_pack4(v.x, v.y, v.z, v.w) << SE_CS_SHIFT;  // id * SE_CS stays in its byte
#endif
}

// Four LDS bytes as HBM ids.
__device__ __forceinline__ uint4 se_unpack_quad(u32 p) {
#if SE_TBL == 2
    p = se_rot4(p, 3u);  // ids (code + 3) mod SE_NMAT
#else
    p = (p >> SE_CS_SHIFT) & (0x01010101u * (0xffu >> SE_CS_SHIFT));
#endif
    return make_uint4(p & 0xffu, (p >> 8) & 0xffu, (p >> 16) & 0xffu, p >> 24);
}

// Block-origin range along one axis at sub-step s: origins o (local) with parity
// `par`, whose block touches the world (g0 + o in [-1, n-1]) and lies in the valid
// region.  A side where the LDS tile cuts into the world loses one cell of
// validity per sub-step; a side where the tile holds the world's WALL border
// does not (those cells never change).
__device__ __forceinline__ void se_block_range(int s, int par, int g0, int L, int n, int& lo, int& cnt) {
    const bool open_lo = g0 > -1, open_hi = g0 + L < n + 1;
    int a = open_lo ? s : 0;
    const int wa = -1 - g0;
    if (wa > a) a = wa;
    if ((a - par) & 1) a++;
    int b = (open_hi ? L - s : L) - 2;
    const int wb = n - 1 - g0;
    if (wb < b) b = wb;
    lo = a;
    cnt = b >= a ? (b - a) / 2 + 1 : 0;
}

// acc + tuple index c0 + N*c1 + N^2*c2 + N^3*c3 of a packed block t = c0 | c1<<8
// | c2<<16 | c3<<24 (bytes < 256, weights for N <= 255): byte dot products with
// constant weights.  Linear in the bytes: over codes k * id it gives k * tuple.
template <u32 N>
__device__ __forceinline__ u32 se_tuple(u32 t, u32 acc = 0u) {
    if (N <= 6) return __builtin_amdgcn_udot4(t, 1u | (N << 8) | ((N * N) << 16) | ((N * N * N) << 24), acc, false);
    if (N <= 15)
        return __builtin_amdgcn_udot4(t, 1u | (N << 8) | ((N * N) << 16), __umul24(t >> 24, N * N * N) + acc, false);
    const u32 lo = __builtin_amdgcn_udot4(t, 1u | (N << 8), acc, false);
    const u32 hi = __builtin_amdgcn_udot4(t, (1u << 16) | (N << 24), 0u, false);
    return lo + __umul24(hi, N * N);
}

// A block's four cells from LDS as c0 | c1<<8 | c2<<16 | c3<<24: two 16-bit
// accesses at even x, four byte accesses at odd x.
template <int LW, bool ODD>
__device__ __forceinline__ u32 se_blk_read(const unsigned char* p0) {
    if (ODD) return se_pack4(p0[0], p0[1], p0[LW], p0[LW + 1]);
    return (u32)*reinterpret_cast<const unsigned short*>(p0) |
           ((u32)*reinterpret_cast<const unsigned short*>(p0 + LW) << 16);
}

template <int LW, bool ODD>
__device__ __forceinline__ void se_blk_write(unsigned char* p0, u32 e) {
    if (ODD) {
        // bytes 0 and 2 from e, bytes 1 and 3 from e >> 8 (byte stores write
        // bits 7:0 or 23:16 of a register)
        u32 e8 = e >> 8;
        __asm__("" : "+v"(e8));
        p0[0] = (unsigned char)e;
        p0[1] = (unsigned char)e8;
        p0[LW] = (unsigned char)(e >> 16);
        p0[LW + 1] = (unsigned char)(e8 >> 16);
    } else {
        *reinterpret_cast<unsigned short*>(p0) = (unsigned short)e;
        *reinterpret_cast<unsigned short*>(p0 + LW) = (unsigned short)(e >> 16);
    }
}

// Odd x offset in an interior tile with LDS rows of 128 bytes: a block's cell
// pairs start at odd columns, so instead of four byte reads each lane reads the
// 2-byte aligned pairs at p0 - 1 of both rows and takes the next lane's pairs
// by a DPP wave shift (lane i gets lane i + 1), and writes the aligned pairs at
// p0 - 1 from its own new block and the previous lane's (wave shift the other
// way): [prev c1, c0] and [prev c3, c2].  The pair left of lane 0's block and
// the one right of the last block lie outside the sub-step's block range, in
// cells no later sub-step of the launch reads (the range shrinks on every side
// of an interior tile), so what lands there does not matter.
#ifndef SE_LDS_DPP
#define SE_LDS_DPP 1
#endif
// a compile-time flag passed to a generic lambda
template <bool B>
struct se_const_bool {
    static constexpr bool value = B;
};
template <int LW>
__device__ __forceinline__ u32 se_blk_read_odd_dpp(const unsigned char* p0) {
    const u32 own = (u32)*reinterpret_cast<const unsigned short*>(p0 - 1) |
                    ((u32)*reinterpret_cast<const unsigned short*>(p0 - 1 + LW) << 16);
    const u32 nb = (u32)__builtin_amdgcn_mov_dpp((int)own, 0x130, 0xf, 0xf, true);  // wave_shl:1, 0 past the wave
    return __builtin_amdgcn_perm(nb, own, 0x06030401u);
}

template <int LW>
__device__ __forceinline__ void se_blk_write_odd_dpp(unsigned char* p0, u32 e) {
    const u32 pe = (u32)__builtin_amdgcn_mov_dpp((int)e, 0x138, 0xf, 0xf, true);  // wave_shr:1
    const u32 w = __builtin_amdgcn_perm(e, pe, 0x06030401u);
    *reinterpret_cast<unsigned short*>(p0 - 1) = (unsigned short)w;
    *reinterpret_cast<unsigned short*>(p0 - 1 + LW) = (unsigned short)(w >> 16);
}

// In an edge tile, the bytes of cells outside the world (columns by colkeep,
// rows from gy) are replaced by WALL, which those cells hold in LDS.
__device__ __forceinline__ u32 se_keep_wall(u32 e, u32 colkeep, int gy, int H) {
    const u32 rowkeep = ((gy >= 0 && gy < H) ? 0x0000ffffu : 0u) | ((gy + 1 >= 0 && gy + 1 < H) ? 0xffff0000u : 0u);
    const u32 keep = colkeep & rowkeep;
    return (e & keep) | ((SE_LWALL * 0x01010101u) & ~keep);
}

__device__ __forceinline__ void se_hash_lanes(u32* h, const u32* xa, const u32* kf, int gy) {
#pragma unroll
    for (int l = 0; l < 4; l++) h[l] = se_hash_used(l) ? se_lowbias(xa[l] + (u32)gy * se_hash_kb(l) + kf[l]) : 0u;
}

// Table path: new packed blocks for two packed blocks ta (world row gya) and tb
// (row gyb): one table read each, at the tuple's entry for the block's hash
// class.  Hashing every block and reading once measured faster than reading a
// hash-independence flag first and hashing only where needed (a second
// dependent LDS read and a branch per block).  xa[l] + gy * kb(l) + kf[l] is the
// hash input of lane l.
// Does a packed block or quad hold a code >= D (NULL or WALL with the coded table)?
__device__ __forceinline__ bool se_has_fixed(u32 t) {
#if SE_TBL == 2
    return ((t + (0x80u - SE_CT_D) * 0x01010101u) & 0x80808080u) != 0u;
#else
    (void)t;
    return false;
#endif
}

#if SE_TBL == 2
// Coded table: a block of codes t with hash lanes h.  Blocks holding NULL or WALL
// (a code >= D) run the generated chain; FIXED = false where the tile holds
// neither.  Otherwise the mirror (lane 0) permutes
// the tuple before the lookup; the tuple's u8 entry names its outcome vector v,
// and entry (v, class r of lanes 1..3, mirror) of the outcome table is a byte
// selector over the block's codes and constant codes: one v_perm builds the new
// block.
template <bool FIXED>
__device__ __forceinline__ u32 se_ct_eval(const u32* __restrict__ lt, u32 t, const u32* h) {
    if (FIXED && __builtin_expect(se_has_fixed(t), 0)) {
        u32 c0 = se_dec(t & 0xffu), c1 = se_dec((t >> 8) & 0xffu), c2 = se_dec((t >> 16) & 0xffu), c3 = se_dec(t >> 24);
        se_block_h(c0, c1, c2, c3, h[0], h[1], h[2], h[3]);
        return se_pack4(se_enc_ok(c0), se_enc_ok(c1), se_enc_ok(c2), se_enc_ok(c3));
    }
    const u32 mirror = h[0] <= SE_CT_MIRROR_T ? 1u : 0u;
    const u32 tm = mirror ? __builtin_amdgcn_perm(t, t, 0x02030001u) : t;
    const u32 v = reinterpret_cast<const unsigned char*>(lt)[se_tuple<SE_CT_D>(tm)];
    // byte offset of entry ((v * R + r) * 2 + mirror) * 8: one 24-bit multiply-add
    const u32 off = __umul24(v, SE_CT_R * 16u) + (se_rclass_ct(h[1], h[2], h[3]) * 16u + (mirror ? 8u : 0u));
    const uint2 o = *reinterpret_cast<const uint2*>(reinterpret_cast<const unsigned char*>(lt + SE_CT_T1_WORDS) + off);
    return __builtin_amdgcn_perm(o.y, t, o.x);
}
#endif

#if SE_TBL == 1
// Full table at odd x, where a block's cells come from LDS as four bytes: the
// entry's byte offset tuple * 4R + 4r as 24-bit multiply-adds of the bytes
// (no packing of the bytes into a word first).
// a * b + c for a, b < 2^24 as one v_mad_u32_u24 (kept as a chain: the
// compiler would otherwise split such sums into shifts, multiplies and adds)
__device__ __forceinline__ u32 se_mad24(u32 a, u32 b, u32 c) {
    u32 d;
    __asm__("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
    return d;
}

template <int LW>
__device__ __forceinline__ u32 se_tbl_at_bytes(const u32* __restrict__ ltbl, const unsigned char* p0, u32 r) {
    constexpr u32 S = SE_TBL_R * 4u / SE_CS, N = SE_NMAT;  // codes carry the factor SE_CS
    u32 off = se_mad24(p0[0], S, r * 4u);
    off = se_mad24(p0[1], S * N, off);
    off = se_mad24(p0[LW], S * N * N, off);
    off = se_mad24(p0[LW + 1], S * N * N * N, off);
    return *reinterpret_cast<const u32*>(reinterpret_cast<const unsigned char*>(ltbl) + off);
}
#endif

// skipa / skipb: the block is left to another pass (its result is not used).
template <bool FIXED>
__device__ __forceinline__ void se_blk_eval2(const u32* __restrict__ ltbl, u32 ta, u32 tb, const u32* xa,
                                             const u32* kf, int gya, int gyb, bool skipa, bool skipb, u32& ea,
                                             u32& eb) {
#if SE_TBL == 1
    u32 ha[4], hb[4];
    se_hash_lanes(ha, xa, kf, gya);
    se_hash_lanes(hb, xa, kf, gyb);
    // byte offsets tuple * 4R + 4r (the class term as one select of constants,
    // added by the tuple's dot product where the codes carry the factor 4R)
    const unsigned char* lb = reinterpret_cast<const unsigned char*>(ltbl);
    constexpr u32 TS = SE_TBL_R * 4u / SE_CS;
    const u32 ra = se_rclass(ha[0], ha[1], ha[2], ha[3]) * 4u, rb = se_rclass(hb[0], hb[1], hb[2], hb[3]) * 4u;
    ea = *reinterpret_cast<const u32*>(lb + (TS == 1u ? se_tuple<SE_NMAT>(ta, ra) : se_tuple<SE_NMAT>(ta) * TS + ra));
    eb = *reinterpret_cast<const u32*>(lb + (TS == 1u ? se_tuple<SE_NMAT>(tb, rb) : se_tuple<SE_NMAT>(tb) * TS + rb));
#elif SE_TBL == 2
    u32 ha[4], hb[4];
    se_hash_lanes(ha, xa, kf, gya);
    se_hash_lanes(hb, xa, kf, gyb);
    ea = se_ct_eval<FIXED>(ltbl, skipa ? 0u : ta, ha);  // a skipped block reads a valid entry
    eb = se_ct_eval<FIXED>(ltbl, skipb ? 0u : tb, hb);
#else
    (void)ltbl; (void)ta; (void)tb; (void)xa; (void)kf; (void)gya; (void)gyb;
    ea = eb = 0u;
#endif
#if SE_TBL != 2
    (void)skipa; (void)skipb;
#endif
}

// Chain path: the generated rule chain for one packed block (all-EMPTY blocks
// skip it).
__device__ __forceinline__ u32 se_blk_eval1(u32 t, const u32* xa, const u32* kf, int gy) {
    if (t == 0u) return 0u;
    u32 h[4];
    se_hash_lanes(h, xa, kf, gy);
    u32 c0 = t & 0xffu, c1 = (t >> 8) & 0xffu, c2 = (t >> 16) & 0xffu, c3 = t >> 24;
    se_block_h(c0, c1, c2, c3, h[0], h[1], h[2], h[3]);
    return se_pack4(c0, c1, c2, c3);
}

// One Margolus sub-step over the block region (fx + 2*bx, fy + 2*by), bx < nbx,
// by < nby, of an LDS tile with row length LW, without modifications.  Lanes of
// a wavefront take consecutive blocks of one block row (64 per pass), the
// wavefronts of the workgroup take every NW-th block row, so each lane keeps its
// column terms (LDS column, the hash's x term) and the row terms are uniform.
// With x offset 0 a block's cell pairs are 2-byte aligned and move as 16-bit
// LDS accesses; with x offset 1 they move as bytes (misaligned 16-bit LDS
// accesses are much slower).  In a tile that reaches past the world (EDGE),
// cells outside the world stay WALL: the result bytes of those cells are
// replaced by WALL before the store.  FIXED: the tile may hold NULL or WALL
// codes, which the coded table does not cover.
#if SE_TBL == 2
// Coded table, edge tiles: the blocks that touch the world border hold WALL
// cells, which the table does not cover.  The main pass leaves them alone and
// the workgroup evaluates them here with the generated chain, one block per
// thread, so the chain does not hold up a whole wavefront for the one border
// lane of every block row.  They are the first/last block row and column of
// the sub-step's block range (blocks partition the cells, so this pass and the
// main pass write disjoint bytes).
template <int LW, int NT, bool ODD>
__device__ __forceinline__ void se_lds_border(unsigned char* lds, int fx, int nbx, int fy, int nby, int gx0, int gy0,
                                              int W, int H, u32 f) {
    if (nbx <= 0 || nby <= 0) return;
    const int gxl = gx0 + fx, gxr = gxl + 2 * (nbx - 1), gyt = gy0 + fy, gyb = gyt + 2 * (nby - 1);
    const bool bl = gxl < 0, br = gxr + 1 >= W && !(bl && nbx == 1);
    const bool bt = gyt < 0, bb = gyb + 1 >= H && !(bt && nby == 1);
    const int nrow = (bt ? 1 : 0) + (bb ? 1 : 0), mid = nby - nrow, ncol = (bl ? 1 : 0) + (br ? 1 : 0);
    const int n = nrow * nbx + ncol * mid;
    for (int i = (int)threadIdx.x; i < n; i += NT) {
        int bx, by;
        if (i < nrow * nbx) {
            const int r = i / nbx;
            bx = i - r * nbx;
            by = (r == 0 && bt) ? 0 : nby - 1;
        } else {
            const int j = i - nrow * nbx, c = j / mid;
            by = (bt ? 1 : 0) + (j - c * mid);
            bx = (c == 0 && bl) ? 0 : nbx - 1;
        }
        const int lx = fx + 2 * bx, ly = fy + 2 * by;
        const int gx = gx0 + lx, gy = gy0 + ly;
        unsigned char* p = lds + ly * LW + lx;
        const u32 t = se_blk_read<LW, ODD>(p);
        u32 c0 = se_dec(t & 0xffu), c1 = se_dec((t >> 8) & 0xffu), c2 = se_dec((t >> 16) & 0xffu), c3 = se_dec(t >> 24);
        se_block(c0, c1, c2, c3, (u32)gx, (u32)gy, f);
        const u32 colkeep =
            ((gx >= 0 && gx < W) ? 0x00ff00ffu : 0u) | ((gx + 1 >= 0 && gx + 1 < W) ? 0xff00ff00u : 0u);
        se_blk_write<LW, ODD>(p, se_keep_wall(se_pack4(se_enc_ok(c0), se_enc_ok(c1), se_enc_ok(c2), se_enc_ok(c3)),
                                              colkeep, gy, H));
    }
}
#endif

template <int LW, int NT, bool ODD, bool EDGE, bool FIXED, bool OPEN = true>
__device__ __forceinline__ void se_lds_substep_t(unsigned char* lds, const u32* __restrict__ ltbl, int fx, int nbx,
                                                 int fy, int nby, int gx0, int gy0, int W, int H, u32 f) {
    constexpr int NW = NT / 64;
    const int lane = (int)(threadIdx.x & 63);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    u32 kf[4];
#pragma unroll
    for (int l = 0; l < 4; l++) kf[l] = f * se_hash_kf(l);
    // odd x in an interior tile (OPEN: the block ranges shrink on every side)
    // with 128-byte LDS rows: aligned pair accesses and DPP shifts instead of
    // byte accesses (se_blk_read_odd_dpp); the lane after the last block stays to
    // write that block's right cells (its own block is not evaluated: t = 0)
    constexpr bool DPP = SE_LDS_DPP && ODD && !EDGE && OPEN && LW == 128;
    for (int bx0 = 0; bx0 < nbx; bx0 += 64) {
        const int bx = bx0 + lane;
        if (DPP ? bx > nbx : bx >= nbx) break;
        const bool valid = bx < nbx;
        const int lx = fx + 2 * bx;
        const int gx = gx0 + lx;
        u32 xa[4];
#pragma unroll
        for (int l = 0; l < 4; l++) xa[l] = (u32)gx * se_hash_ka(l);
        u32 colkeep = 0xffffffffu;
        if (EDGE)
            colkeep = ((gx >= 0 && gx < W) ? 0x00ff00ffu : 0u) | ((gx + 1 >= 0 && gx + 1 < W) ? 0xff00ff00u : 0u);
        // coded table, edge tile: blocks touching the world border (WALL cells,
        // which the table does not cover) are left to se_lds_border
        constexpr bool BORDER = SE_TBL == 2 && EDGE;
        const bool colb = BORDER && (gx < 0 || gx + 1 >= W);
        // full table: two block rows (by, by + NW) per iteration, their reads,
        // lookups and writes interleaved so the two LDS chains overlap; the coded
        // table and the branchy rule chain run one block row per iteration (each
        // measured faster that way).  Pairs of rows run in one loop, a last single
        // row after it, so the second row's LDS offset is an immediate.
        constexpr int RPI = SE_TBL == 1 ? 2 : 1;
        // hash inputs of lane l for the current block row (x*ka + y*kb + frame*kf),
        // stepped with the row; the second row's are hx + 2*NW*kb
        u32 hx[4];
#pragma unroll
        for (int l = 0; l < 4; l++) hx[l] = xa[l] + (u32)(gy0 + fy + 2 * wv) * se_hash_kb(l) + kf[l];
        const u32 kz[4] = {0u, 0u, 0u, 0u};
        auto row = [&](int by, auto tag) {
            constexpr bool two = decltype(tag)::value;
            const int gya = gy0 + fy + 2 * by, gyb = gya + 2 * NW;
            unsigned char* pa = lds + (fy + 2 * by) * LW + lx;
            unsigned char* pb = pa + 2 * NW * LW;  // read only for a pair of rows
            const bool skipa = colb || (BORDER && (gya < 0 || gya + 1 >= H));
            const bool skipb = colb || (BORDER && (gyb < 0 || gyb + 1 >= H));
            u32 ea, eb;
            if (DPP) {
                u32 ta = se_blk_read_odd_dpp<LW>(pa), tb = two ? se_blk_read_odd_dpp<LW>(pb) : 0u;
                ta = valid ? ta : 0u;
                tb = valid ? tb : 0u;
#if SE_TBL == 0
                eb = 0u;
                ea = se_blk_eval1(ta, hx, kz, 0);
#else
                se_blk_eval2<FIXED>(ltbl, ta, tb, hx, kz, 0, 2 * NW, false, false, ea, eb);
#endif
                se_blk_write_odd_dpp<LW>(pa, ea);
                if (two) se_blk_write_odd_dpp<LW>(pb, eb);
                return;
            }
#if SE_TBL == 1
            if (ODD) {  // full table at odd x: offsets from the cell bytes
                u32 ha[4], hb[4];
                se_hash_lanes(ha, hx, kz, 0);
                se_hash_lanes(hb, hx, kz, 2 * NW);
                ea = se_tbl_at_bytes<LW>(ltbl, pa, se_rclass(ha[0], ha[1], ha[2], ha[3]));
                eb = two ? se_tbl_at_bytes<LW>(ltbl, pb, se_rclass(hb[0], hb[1], hb[2], hb[3])) : 0u;
            } else {
                const u32 ta = se_blk_read<LW, ODD>(pa), tb = two ? se_blk_read<LW, ODD>(pb) : 0u;
                se_blk_eval2<FIXED>(ltbl, ta, tb, hx, kz, 0, 2 * NW, skipa, skipb, ea, eb);
            }
#elif SE_TBL == 2
            const u32 ta = se_blk_read<LW, ODD>(pa);
            se_blk_eval2<FIXED>(ltbl, ta, 0u, hx, kz, 0, 2 * NW, skipa, skipb, ea, eb);
#else
            eb = 0u;
            ea = se_blk_eval1(se_blk_read<LW, ODD>(pa), hx, kz, 0);
#endif
            if (EDGE && !BORDER) {
                ea = se_keep_wall(ea, colkeep, gya, H);
                eb = se_keep_wall(eb, colkeep, gyb, H);
            }
            if (!skipa) se_blk_write<LW, ODD>(pa, ea);
            if (two && !skipb) se_blk_write<LW, ODD>(pb, eb);
        };
        int by = wv;
        if (RPI == 2) {
            for (; by + NW < nby; by += 2 * NW) {
                row(by, se_const_bool<true>());
#pragma unroll
                for (int l = 0; l < 4; l++) {
                    hx[l] += (u32)(4 * NW) * se_hash_kb(l);
                    // one register per lane, stepped by a constant (the compiler
                    // would otherwise rebuild it from two parts every row)
                    if (se_hash_used(l)) __asm__("" : "+v"(hx[l]));
                }
            }
        }
        for (; by < nby; by += NW) {
            row(by, se_const_bool<false>());
#pragma unroll
            for (int l = 0; l < 4; l++) hx[l] += (u32)(2 * NW) * se_hash_kb(l);
        }
    }
#if SE_TBL == 2
    if (EDGE) se_lds_border<LW, NT, ODD>(lds, fx, nbx, fy, nby, gx0, gy0, W, H, f);
#endif
}

// One Margolus sub-step without modifications; odd x offsets, edge tiles and
// (coded table) interior tiles free of NULL and WALL get their own
// instantiations.  edge: the tile needs the edge-tile block loop; open: the
// tile lies inside the world (a full table that keeps WALL runs edge tiles
// through the interior loop, but not through its aligned-pair odd path, which
// writes cells just outside the block range: in an edge tile those can be the
// world's WALL border).
template <int LW, int NT>
__device__ __forceinline__ void se_lds_substep(unsigned char* lds, const u32* __restrict__ ltbl, int fx, int nbx,
                                               int fy, int nby, int gx0, int gy0, int W, int H, u32 f, bool edge,
                                               bool fixed, bool open) {
    const bool odd = (fx & 1) != 0;  // block origins share the parity of the x offset
    const bool clean = SE_TBL == 2 && !edge && !fixed;
    if (SE_TBL == 2 && edge && !fixed) {  // edge tile with no NULL or WALL inside the world
        if (odd) se_lds_substep_t<LW, NT, true, true, false>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        else se_lds_substep_t<LW, NT, false, true, false>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        return;
    }
    if (odd) {
        if (edge) se_lds_substep_t<LW, NT, true, true, true>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        else if (SE_TBL == 1 && !open)
            se_lds_substep_t<LW, NT, true, false, true, false>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        else if (clean) se_lds_substep_t<LW, NT, true, false, false>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        else se_lds_substep_t<LW, NT, true, false, true>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
    } else {
        if (edge) se_lds_substep_t<LW, NT, false, true, true>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        else if (clean) se_lds_substep_t<LW, NT, false, false, false>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
        else se_lds_substep_t<LW, NT, false, false, true>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f);
    }
}

template <int TW, int TH, int HB, int NT>
__device__ __forceinline__ void se_lds_body(const u32* __restrict__ in, u32* __restrict__ out, int W, int H,
                                            int pitch, int frame0, int nsub, int tiles_x, int tile,
                                            const se_mod* __restrict__ mods, int nmods,
                                            const u32* __restrict__ ltbl, unsigned char* lds, u32* flag, u32 seq,
                                            unsigned char* __restrict__ sleep_out) {
    constexpr int LW = TW + 2 * HB, LH = TH + 2 * HB, LQ = LW / 4;
    // The thread index passes through an empty asm so that per-thread offsets of
    // the tile load/store are recomputed per tile: hoisted out of a persistent
    // workgroup's tile loop they would stay live in ~80 VGPRs and cut occupancy.
    int tid = (int)threadIdx.x;
    __asm__ volatile("" : "+v"(tid));
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int gx0 = tx * TW - HB, gy0 = ty * TH - HB;
    const bool interior = gx0 >= 0 && gy0 >= 0 && gx0 + LW <= W && gy0 + LH <= H;
    u32* lds32 = reinterpret_cast<u32*>(lds);
    bool fixed = false;  // this thread stored a NULL or WALL code (coded table)
    // interior tiles: the code of the tile's first cell in all four bytes; quads
    // that differ from it make the tile unsettled
    const u32 ref = interior ? se_enc(in[(size_t)gy0 * pitch + gx0]) * 0x01010101u : 0u;
    bool mixed = false;
    // whole tile inside the world, workgroup a whole number of tile rows wide:
    // each thread keeps one quad column and steps down the rows by pointer
    // increments (no per-quad index arithmetic), U loads in flight before any is
    // decoded into LDS (the whole tile at once up to 16 per thread: +1.5 % over 8)
    constexpr int DROW = NT % LQ == 0 ? NT / LQ : 1, KMAX = LH / DROW;
    constexpr bool ROWS = NT % LQ == 0 && LH % DROW == 0 && KMAX % 8 == 0;
    if (ROWS && interior) {
        constexpr int U = KMAX <= 16 ? KMAX : 8;
        const int lq = tid % LQ, ly0 = tid / LQ;
        const u32* src = in + (size_t)(gy0 + ly0) * pitch + gx0 + 4 * lq;
        const size_t sstep = (size_t)DROW * pitch;
        u32* dst = lds32 + ly0 * LQ + lq;
#pragma unroll
        for (int k0 = 0; k0 < KMAX; k0 += U) {
            uint4 v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                v[u] = *reinterpret_cast<const uint4*>(src);
                src += sstep;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const u32 q = se_pack_quad(v[u]);
                fixed |= se_has_fixed(q);
                mixed |= q != ref;
                dst[(k0 + u) * DROW * LQ] = q;
            }
        }
    } else if (interior) {
        // whole tile inside the world: every thread issues U quad loads before
        // any of them is decoded into LDS, so a thread waits U loads at a time
        constexpr int TOTAL = LQ * LH, U = 8;
        for (int i0 = 0; i0 < TOTAL; i0 += U * NT) {
            uint4 v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = min(i0 + u * NT + tid, TOTAL - 1);  // clamped: unconditional loads
                const int ly = i / LQ, lq = i - ly * LQ;
                v[u] = *reinterpret_cast<const uint4*>(in + (size_t)(gy0 + ly) * pitch + gx0 + 4 * lq);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * NT + tid;
                if (i < TOTAL) {
                    const u32 q = se_pack_quad(v[u]);
                    fixed |= se_has_fixed(q);
                    mixed |= q != ref;
                    lds32[i] = q;
                }
            }
        }
    } else if (ROWS) {
        // edge tile, rows mapping: the quads that touch the world or its one-cell
        // WALL border.  U loads are in flight at a time from addresses clamped into
        // the world's rows (always valid; rows are at least 4 cells of pitch); a
        // quad wholly inside the world takes the loaded value, a quad across the
        // world's side is read cell by cell, a row outside the world is WALL.
        constexpr int U = 8;
        const int ly_lo = max(0, -1 - gy0), ly_hi = min(LH, H + 1 - gy0);
        const int lq_lo = max(0, (-1 - gx0) >> 2), lq_hi = min(LQ, (W + 1 - gx0 + 3) >> 2);
        const int lq = tid % LQ, ly0 = tid / LQ;
        const int gx = gx0 + 4 * lq;
        const bool qin = lq >= lq_lo && lq < lq_hi;
        const bool xfull = gx >= 0 && gx + 3 < W;
        const int gxc = xfull ? gx : 0;
#pragma unroll
        for (int k0 = 0; k0 < KMAX; k0 += U) {
            uint4 v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int gy = gy0 + ly0 + (k0 + u) * DROW;
                const int gyc = min(max(gy, 0), H - 1);
                v[u] = *reinterpret_cast<const uint4*>(in + (size_t)gyc * pitch + gxc);
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int ly = ly0 + (k0 + u) * DROW, gy = gy0 + ly;
                if (qin && ly >= ly_lo && ly < ly_hi) {
                    u32 q = SE_LWALL * 0x01010101u;
                    if (gy >= 0 && gy < H) {
                        if (xfull) {
                            q = se_pack_quad(v[u]);
                            fixed |= se_has_fixed(q);
                        } else {
                            const u32* row = in + (size_t)gy * pitch;
                            u32 c[4];
#pragma unroll
                            for (int j = 0; j < 4; j++) {
                                const bool inw = gx + j >= 0 && gx + j < W;
                                c[j] = inw ? se_enc(row[gx + j]) : SE_LWALL;
                                fixed |= inw && se_has_fixed(c[j]);
                            }
                            q = se_pack4(c[0], c[1], c[2], c[3]);
                        }
                    }
                    lds32[ly * LQ + lq] = q;
                }
            }
        }
    } else {
        // tile load: rows/quads that touch the world plus its one-cell WALL border
        const int ly_lo = max(0, -1 - gy0), ly_hi = min(LH, H + 1 - gy0);
        const int lq_lo = max(0, (-1 - gx0) >> 2), lq_hi = min(LQ, (W + 1 - gx0 + 3) >> 2);
        const int nq = max(0, lq_hi - lq_lo), nr = max(0, ly_hi - ly_lo);
        for (int i = tid; i < nq * nr; i += NT) {
            const int ry = i / nq;
            const int ly = ly_lo + ry, lq = lq_lo + (i - ry * nq);
            const int gx = gx0 + 4 * lq, gy = gy0 + ly;
            u32 v0 = SE_LWALL, v1 = SE_LWALL, v2 = SE_LWALL, v3 = SE_LWALL;
            if (gy >= 0 && gy < H) {
                const u32* row = in + (size_t)gy * pitch;
                if (gx >= 0 && gx + 3 < W) {
                    const uint4 v = *reinterpret_cast<const uint4*>(row + gx);
                    v0 = se_enc(v.x); v1 = se_enc(v.y); v2 = se_enc(v.z); v3 = se_enc(v.w);
                    fixed |= se_has_fixed(se_pack4(v0, v1, v2, v3));
                } else {
                    if (gx >= 0 && gx < W) { v0 = se_enc(row[gx]); fixed |= se_has_fixed(v0); }
                    if (gx + 1 >= 0 && gx + 1 < W) { v1 = se_enc(row[gx + 1]); fixed |= se_has_fixed(v1); }
                    if (gx + 2 >= 0 && gx + 2 < W) { v2 = se_enc(row[gx + 2]); fixed |= se_has_fixed(v2); }
                    if (gx + 3 >= 0 && gx + 3 < W) { v3 = se_enc(row[gx + 3]); fixed |= se_has_fixed(v3); }
                }
            }
            lds32[ly * LQ + lq] = se_pack4(v0, v1, v2, v3);
        }
    }
    // an interior tile runs the table without the NULL/WALL test when no cell
    // holds either (table results never introduce them).  Threads that stored
    // such a cell write this item's sequence number into the LDS flag word; an
    // older number means clean, so the word needs no clearing between items.
    if (SE_TBL == 2 && fixed) flag[0] = seq;
    if (mixed) flag[1] = seq;
    __syncthreads();
    bool tile_fixed = SE_TBL == 2 && flag[0] == seq;
    // A settled tile: interior, every cell (halo included) the same code c with
    // bit c of se_stable_mask set, and no modification near it.  Every block of
    // every sub-step is then four cells of c, which stay as they are: the
    // sub-steps are skipped and the tile is stored as loaded.
    // (se_stable_mask is indexed by code with the coded table, by id otherwise)
    const u32 cref = ref & 0xffu, cbit = SE_TBL == 2 ? cref : se_dec(cref);
    const bool settled = interior && flag[1] != seq && cbit < 32u && ((se_stable_mask >> cbit) & 1u) != 0u &&
                         (nmods == 0 || !se_mods_touch(mods, nmods, gx0, gx0 + LW - 1, gy0, gy0 + LH - 1));
    if (sleep_out != nullptr && tid == 0) sleep_out[tile] = settled ? (unsigned char)cref : (unsigned char)0xffu;
    // edge tiles need the WALL restoring of their block loop unless the full
    // table keeps WALL cells by itself
#if SE_TBL == 1
    const bool mask_wall = !interior && se_wall_kept == 0u;
#else
    const bool mask_wall = !interior;
#endif
    for (int s = 0; s < (settled ? 0 : nsub); s++) {
        const u32 f = (u32)(frame0 + s);
        int ox, oy;
        se_offset(f, ox, oy);
        int fx, nbx, fy, nby;
        // with fewer sub-steps than the halo, the outer HB - nsub rings cannot
        // reach the interior: the ranges shrink as in a tile of halo nsub
        const int sr = s + (nsub < HB ? HB - nsub : 0);
        se_block_range(sr, ox, gx0, LW, W, fx, nbx);
        se_block_range(sr, oy, gy0, LH, H, fy, nby);
        const int nb = nbx * nby;
        const bool use_mods = s == 0 && nmods > 0;
        if (!use_mods) {
            se_lds_substep<LW, NT>(lds, ltbl, fx, nbx, fy, nby, gx0, gy0, W, H, f, mask_wall, tile_fixed, interior);
        } else if (nb > 0) {
            const int dby = NT / nbx, dbx = NT - dby * nbx;
            int by = tid / nbx, bx = tid - by * nbx;
            for (int b = tid; b < nb; b += NT) {
                const int lx = fx + 2 * bx, ly = fy + 2 * by;
                bx += dbx;
                by += dby;
                if (bx >= nbx) { bx -= nbx; by += 1; }
                unsigned char* p0 = lds + ly * LW + lx;
                unsigned char* p1 = p0 + LW;
                u32 c0, c1, c2, c3;
                if (ox == 0) {  // lx even: 2-byte aligned pairs
                    const u32 u = *reinterpret_cast<const unsigned short*>(p0);
                    const u32 v = *reinterpret_cast<const unsigned short*>(p1);
                    c0 = u & 0xffu; c1 = u >> 8; c2 = v & 0xffu; c3 = v >> 8;
                } else {
                    c0 = p0[0]; c1 = p0[1]; c2 = p1[0]; c3 = p1[1];
                }
                c0 = se_dec(c0); c1 = se_dec(c1); c2 = se_dec(c2); c3 = se_dec(c3);
                const int gx = gx0 + lx, gy = gy0 + ly;
#if SE_TBL == 1
                const u32 base = (c0 + SE_NMAT * (c1 + SE_NMAT * (c2 + SE_NMAT * c3))) * SE_TBL_R;
                const u32 r = se_rclass(se_hash_lane(0, (u32)gx, (u32)gy, f), se_hash_lane(1, (u32)gx, (u32)gy, f),
                                        se_hash_lane(2, (u32)gx, (u32)gy, f), se_hash_lane(3, (u32)gx, (u32)gy, f));
                const u32 e = ltbl[base + r];  // LDS codes
                c0 = se_dec(e & 0xffu); c1 = se_dec((e >> 8) & 0xffu); c2 = se_dec((e >> 16) & 0xffu);
                c3 = se_dec(e >> 24);
#else
                se_block(c0, c1, c2, c3, (u32)gx, (u32)gy, f);
#endif
                if (use_mods) {
                    int o;
                    if ((o = se_mod_override(mods, nmods, gx, gy)) >= 0) c0 = (u32)o;
                    if ((o = se_mod_override(mods, nmods, gx + 1, gy)) >= 0) c1 = (u32)o;
                    if ((o = se_mod_override(mods, nmods, gx, gy + 1)) >= 0) c2 = (u32)o;
                    if ((o = se_mod_override(mods, nmods, gx + 1, gy + 1)) >= 0) c3 = (u32)o;
                }
                c0 = se_enc_ok(c0); c1 = se_enc_ok(c1); c2 = se_enc_ok(c2); c3 = se_enc_ok(c3);
                if (interior) {
                    fixed |= se_has_fixed(se_pack4(c0, c1, c2, c3));
                    if (ox == 0) {
                        *reinterpret_cast<unsigned short*>(p0) = (unsigned short)(c0 | (c1 << 8));
                        *reinterpret_cast<unsigned short*>(p1) = (unsigned short)(c2 | (c3 << 8));
                    } else {
                        p0[0] = (unsigned char)c0; p0[1] = (unsigned char)c1;
                        p1[0] = (unsigned char)c2; p1[1] = (unsigned char)c3;
                    }
                } else {
                    const bool inx0 = gx >= 0 && gx < W, inx1 = gx + 1 >= 0 && gx + 1 < W;
                    const bool iny0 = gy >= 0 && gy < H, iny1 = gy + 1 >= 0 && gy + 1 < H;
                    if (inx0 && iny0) { p0[0] = (unsigned char)c0; fixed |= se_has_fixed(c0); }
                    if (inx1 && iny0) { p0[1] = (unsigned char)c1; fixed |= se_has_fixed(c1); }
                    if (inx0 && iny1) { p1[0] = (unsigned char)c2; fixed |= se_has_fixed(c2); }
                    if (inx1 && iny1) { p1[1] = (unsigned char)c3; fixed |= se_has_fixed(c3); }
                }
            }
        }
        if (SE_TBL == 2 && use_mods && fixed) flag[0] = seq;  // brushes may place WALL
        __syncthreads();
        if (SE_TBL == 2 && use_mods) tile_fixed = flag[0] == seq;
    }
    // interior store: quads from LDS, 16-byte global stores where inside the world
    constexpr int TQ = TW / 4;
    // rows mapping: threads take quad column tid % QW (QW a power of two >= TQ;
    // columns >= TQ idle) and step down the rows by pointer increments
    constexpr int QW = TQ <= 16 ? 16 : TQ <= 32 ? 32 : TQ <= 64 ? 64 : 128;
    constexpr int DR = NT % QW == 0 ? NT / QW : 1, KS = TH / DR;
    constexpr bool SROWS = NT % QW == 0 && TH % DR == 0;
    if (SROWS && gx0 + HB + TW <= W && gy0 + HB + TH <= H) {
        const int lq = tid % QW, ly0 = tid / QW;
        if (lq < TQ) {
            const u32* s = lds32 + (ly0 + HB) * LQ + HB / 4 + lq;
            u32* d = out + (size_t)(gy0 + HB + ly0) * pitch + gx0 + HB + 4 * lq;
            const size_t dstep = (size_t)DR * pitch;
            constexpr int U = 4;
#pragma unroll
            for (int k0 = 0; k0 < KS; k0 += U) {
                u32 p[U];
#pragma unroll
                for (int u = 0; u < U; u++) p[u] = k0 + u < KS ? s[(k0 + u) * DR * LQ] : 0u;
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (k0 + u < KS) {
                        se_store_quad<true>(d, se_unpack_quad(p[u]));
                        d += dstep;
                    }
                }
            }
        }
        __syncthreads();  // LDS is reused by the next item of a persistent workgroup
        return;
    }
    if (gx0 + HB + TW <= W && gy0 + HB + TH <= H) {
        constexpr int TOTAL = TQ * TH, U = 4;
        for (int i0 = 0; i0 < TOTAL; i0 += U * NT) {
            u32 p[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * NT + tid;
                const int ly = i / TQ, lq = i - ly * TQ;
                p[u] = i < TOTAL ? lds32[((ly + HB) * LW + HB) / 4 + lq] : 0u;
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int i = i0 + u * NT + tid;
                if (i < TOTAL) {
                    const int ly = i / TQ, lq = i - ly * TQ;
                    const uint4 v = se_unpack_quad(p[u]);
                    se_store_quad<true>(out + (size_t)(gy0 + HB + ly) * pitch + gx0 + HB + 4 * lq, v);
                }
            }
        }
        __syncthreads();  // LDS is reused by the next item of a persistent workgroup
        return;
    }
    const int sy_hi = max(0, min(TH, H - (gy0 + HB))), sq_hi = max(0, min(TQ, (W - (gx0 + HB) + 3) >> 2));
    if (SROWS) {  // tile at the world's edge: the rows mapping, cut to the world
        const int lq = tid % QW, ly0 = tid / QW;
        if (lq < sq_hi) {
            const int gx = gx0 + HB + 4 * lq;
            const u32* s = lds32 + (ly0 + HB) * LQ + HB / 4 + lq;
#pragma unroll
            for (int k = 0; k < KS; k++) {
                const int ly = ly0 + k * DR;
                if (ly < sy_hi) {
                    const uint4 v = se_unpack_quad(s[k * DR * LQ]);
                    u32* dst = out + (size_t)(gy0 + HB + ly) * pitch + gx;
                    if (gx + 3 < W) {
                        se_store_quad<true>(dst, v);
                    } else {
                        dst[0] = v.x;
                        if (gx + 1 < W) dst[1] = v.y;
                        if (gx + 2 < W) dst[2] = v.z;
                    }
                }
            }
        }
        __syncthreads();  // LDS is reused by the next item of a persistent workgroup
        return;
    }
    for (int i = tid; i < sq_hi * sy_hi; i += NT) {
        const int ly = i / sq_hi, lq = i - ly * sq_hi;
        const int gx = gx0 + HB + 4 * lq, gy = gy0 + HB + ly;
        const u32 p = lds32[((ly + HB) * LW + HB) / 4 + lq];
        const uint4 v = se_unpack_quad(p);
        u32* dst = out + (size_t)gy * pitch + gx;
        if (gx + 3 < W) {
            se_store_quad<true>(dst, v);
        } else {
            dst[0] = v.x;
            if (gx + 1 < W) dst[1] = v.y;
            if (gx + 2 < W) dst[2] = v.z;
        }
    }
    __syncthreads();  // LDS is reused by the next item of a persistent workgroup
}

// The table (SE_TBL_WORDS u32 words) into LDS after the tile; the two flag
// words after the table (NULL/WALL seen, mixed tile) start at 0.  Full-table
// entries (ids) become LDS codes (id * SE_CS, each in its own byte).
__device__ __forceinline__ u32 se_tbl_code(u32 e) { return SE_TBL == 1 ? e << SE_CS_SHIFT : e; }
__device__ __forceinline__ const u32* se_stage_table(const u32* __restrict__ tbl, unsigned char* lds, int lds_tile,
                                                     int nt, u32* flag) {
    if (threadIdx.x < 2) flag[threadIdx.x] = 0u;
#if SE_TBL
    u32* ltbl = reinterpret_cast<u32*>(lds + ((lds_tile + 15) & ~15));
    // 16-byte loads, U per thread in flight before their LDS stores
    constexpr int tbl_words = (int)SE_TBL_WORDS, Q = tbl_words / 4;
    constexpr int U = 4;
    const uint4* src = reinterpret_cast<const uint4*>(tbl);
    uint4* dst = reinterpret_cast<uint4*>(ltbl);
    for (int i0 = 0; i0 < Q; i0 += U * nt) {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * nt + (int)threadIdx.x;
            v[u] = src[min(i, Q - 1)];  // unconditional (clamped) loads keep v[] in registers
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = i0 + u * nt + (int)threadIdx.x;
            const uint4 c = make_uint4(se_tbl_code(v[u].x), se_tbl_code(v[u].y), se_tbl_code(v[u].z),
                                       se_tbl_code(v[u].w));
            if (i < Q) dst[i] = c;
        }
    }
    for (int i = 4 * Q + (int)threadIdx.x; i < tbl_words; i += nt) ltbl[i] = se_tbl_code(tbl[i]);
    __syncthreads();
    return ltbl;
#else
    (void)tbl; (void)lds; (void)lds_tile; (void)nt;
    __syncthreads();
    return nullptr;
#endif
}

// Per-world record of a batched run (device memory, filled by the host).
struct se_bworld {
    u32* buf[2];     // the world's two grids; buf[cur0] holds the grid at batch start
    int W, H, pitch, tiles_x, ntiles;
    int frame0;      // frame of the world's first step in this run
    int steps;       // steps the world advances in this run
    int cur0;
    const se_mod* mods;  // modifications for the world's first step in this run
    int nmods;
    int pad;
};

// Index of the world owning item `item`: prefix[w] <= item < prefix[w+1].
__device__ __forceinline__ int se_owner(const int* __restrict__ prefix, int n, int item) {
    int lo = 0, hi = n;  // prefix[lo] <= item < prefix[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Workgroup b takes tile (b % 8) * q + b / 8 (q = n / 8; the last n % 8 keep
// their index), so each XCD walks one contiguous run of tiles and neighbouring
// tiles, which share halo lines, are fetched through the same L2.
__device__ __forceinline__ int se_xcd_tile(int b, int n) {
    const int q = n >> 3;
    if (b >= (q << 3)) return b;
    return (b & 7) * q + (b >> 3);
}

// One world's tiles: the contiguous runs of se_xcd_tile taken over the tile
// rows in the order 0, 8, 16, ..., 1, 9, 17, ..., so that each XCD walks rows
// spread over the whole world (row j, j + 8, ...; neighbouring tiles of a row
// stay on one XCD).  Work that is uneven in y -- a settled floor, an empty sky
// -- then spreads evenly over the XCDs instead of leaving some of them idle.
__device__ __forceinline__ int se_world_tile(int b, int n, int tiles_x) {
    const int idx = se_xcd_tile(b, n);
    const int tiles_y = n / tiles_x, p = idx / tiles_x, col = idx - p * tiles_x;
    int first = 0, row = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {  // rows r = j mod 8 take positions [first, first + count)
        const int count = (tiles_y - j + 7) >> 3;
        if (p >= first && p < first + count) row = j + 8 * (p - first);
        first += count;
    }
    return row * tiles_x + col;
}

// Tile sleeping across launches of one world (se_step_lds_*).  sleep_in[t] is
// the code c when tile t was settled in the previous launch (every cell of its
// region, halo included, c with a stable uniform block; no brush near), else
// 0xff.  If t and its eight neighbours were all settled with the same c and no
// brush of this launch comes near t, then:
//   * t's region in the input (t's interior and the neighbours' interiors
//     around it) is still all c, so the sub-steps leave it as it is, and
//   * the output buffer, which held the previous launch's input, already holds c
//     in t's interior;
// so t neither loads nor stores anything: it records c again and returns.
// Interior tiles only (an edge tile is never settled, so neither it nor a tile
// beside it sleeps).
template <int TW, int TH, int HB>
__device__ __forceinline__ bool se_tile_sleeps(const unsigned char* __restrict__ sleep_in,
                                               unsigned char* __restrict__ sleep_out, int tile, int tiles_x,
                                               int ntiles, const se_mod* __restrict__ mods, int nmods) {
    if (sleep_in == nullptr) return false;
    const u32 c = sleep_in[tile];
    if (c == 0xffu) return false;
    const int tx = tile % tiles_x, ty = tile / tiles_x, tiles_y = ntiles / tiles_x;
    if (tx == 0 || ty == 0 || tx == tiles_x - 1 || ty == tiles_y - 1) return false;
    bool same = true;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) same = same && sleep_in[tile + dy * tiles_x + dx] == c;
    if (!same) return false;
    const int gx0 = tx * TW - HB, gy0 = ty * TH - HB;
    if (nmods > 0 && se_mods_touch(mods, nmods, gx0, gx0 + TW + 2 * HB - 1, gy0, gy0 + TH + 2 * HB - 1)) return false;
    if (threadIdx.x == 0) sleep_out[tile] = (unsigned char)c;
    return true;
}

// LDS (static, so every table offset is an immediate): tile bytes rounded to 16,
// the table (SE_TBL_WORDS u32), a 16-byte slot holding the tile's flag words
// (NULL/WALL seen, mixed tile)
//  se_step_lds_*  : one world, one workgroup per tile (workgroups that start as
//                   others retire keep tile loads and tile compute overlapped;
//                   measured faster than persistent workgroups, which move in step).
//  se_batch_lds_* : persistent over many worlds: workgroups stride over the
//                   tiles of worlds [0, n_active) (sorted by steps, longest first)
//                   and advance each by min(k, steps - round*k) sub-steps; a world
//                   that fits one tile runs with k >= all its steps in one item.
#define SE_LDS_KERNEL(TW, TH, HB, NT)                                                                         \
    extern "C" __global__ __launch_bounds__(NT) void se_step_lds_##TW##x##TH##_h##HB##_t##NT(                \
        const u32* __restrict__ in, u32* __restrict__ out, int W, int H, int pitch, int frame0, int nsub,     \
        int tiles_x, const se_mod* __restrict__ mods, int nmods, const u32* __restrict__ tbl,                 \
        const unsigned char* __restrict__ sleep_in, unsigned char* __restrict__ sleep_out) {                  \
        constexpr int LT = ((TW + 2 * HB) * (TH + 2 * HB) + 15) & ~15;                                         \
        __shared__ __align__(16) unsigned char se_lds[LT + 4 * SE_TBL_WORDS + 16];                            \
        const int tile = se_world_tile((int)blockIdx.x, (int)gridDim.x, tiles_x);                             \
        if (se_tile_sleeps<TW, TH, HB>(sleep_in, sleep_out, tile, tiles_x, (int)gridDim.x, mods, nmods)) return; \
        u32* fl = reinterpret_cast<u32*>(se_lds + LT + 4 * SE_TBL_WORDS);                                      \
        const u32* lt = se_stage_table(tbl, se_lds, (TW + 2 * HB) * (TH + 2 * HB), NT, fl);                  \
        se_lds_body<TW, TH, HB, NT>(in, out, W, H, pitch, frame0, nsub, tiles_x, tile, mods, nmods, lt, se_lds, \
                                    fl, 1u, sleep_out);                                                       \
    }                                                                                                         \
    extern "C" __global__ __launch_bounds__(NT) void se_batch_lds_##TW##x##TH##_h##HB##_t##NT(               \
        const se_bworld* __restrict__ worlds, const int* __restrict__ prefix, int n_active, int round, int k, \
        const u32* __restrict__ tbl) {                                                                        \
        constexpr int LT = ((TW + 2 * HB) * (TH + 2 * HB) + 15) & ~15;                                         \
        __shared__ __align__(16) unsigned char se_lds[LT + 4 * SE_TBL_WORDS + 16];                            \
        u32* fl = reinterpret_cast<u32*>(se_lds + LT + 4 * SE_TBL_WORDS);                                      \
        const u32* lt = se_stage_table(tbl, se_lds, (TW + 2 * HB) * (TH + 2 * HB), NT, fl);                  \
        const int n_items = prefix[n_active];                                                                 \
        u32 seq = 0;                                                                                          \
        for (int item = (int)blockIdx.x; item < n_items; item += (int)gridDim.x) {                            \
            const int w = se_owner(prefix, n_active, item);                                                   \
            const se_bworld& B = worlds[w];                                                                   \
            const int done = round * k;                                                                       \
            const int nsub = min(k, B.steps - done);                                                          \
            const int cur = (B.cur0 + round) & 1;                                                             \
            se_lds_body<TW, TH, HB, NT>(B.buf[cur], B.buf[cur ^ 1], B.W, B.H, B.pitch, B.frame0 + done, nsub,  \
                                        B.tiles_x, item - prefix[w], done == 0 ? B.mods : nullptr,            \
                                        done == 0 ? B.nmods : 0, lt, se_lds, fl, ++seq, nullptr);             \
        }                                                                                                     \
    }

// ---------------------------------------------------------------------------
// Seeded i.i.d. fill of the benchmark workloads: r = hi32(splitmix64((seed << 40)
// ^ (y*W + x))); the first k < n-1 with r < thr[k] picks ids[k], else ids[n-1];
// rows y >= H - stone_rows take stone_id.  Padding columns are left alone.
// ---------------------------------------------------------------------------
__device__ __forceinline__ u64 se_splitmix64(u64 z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct se_fillspec {  // fill parameters passed by value
    int n, stone_rows;
    u32 stone_id, pad;
    u32 ids[16], thr[16];
};

extern "C" __global__ __launch_bounds__(256) void se_fill_spec(u32* __restrict__ g, int W, int H, int pitch,
                                                               u64 seed, se_fillspec sp) {
    const size_t total = (size_t)W * (size_t)H;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t y = i / (size_t)W, x = i - y * (size_t)W;
        u32 v;
        if ((long)y >= (long)H - sp.stone_rows) {
            v = sp.stone_id;
        } else {
            const u32 r = (u32)(se_splitmix64((seed << 40) ^ (u64)i) >> 32);
            v = sp.ids[sp.n - 1];
            for (int k = 0; k < sp.n - 1; k++)
                if (r < sp.thr[k]) { v = sp.ids[k]; break; }
        }
        g[y * (size_t)pitch + x] = v;
    }
}

extern "C" __global__ __launch_bounds__(256) void se_fill_seeded(u32* __restrict__ g, int W, int H, int pitch,
                                                                 u64 seed, const u32* __restrict__ ids,
                                                                 const u32* __restrict__ thr, int n,
                                                                 int stone_rows, u32 stone_id) {
    const size_t total = (size_t)W * (size_t)H;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
        const size_t y = i / (size_t)W, x = i - y * (size_t)W;
        u32 v;
        if ((long)y >= (long)H - stone_rows) {
            v = stone_id;
        } else {
            const u32 r = (u32)(se_splitmix64((seed << 40) ^ (u64)i) >> 32);
            v = ids[n - 1];
            for (int k = 0; k < n - 1; k++)
                if (r < thr[k]) { v = ids[k]; break; }
        }
        g[y * (size_t)pitch + x] = v;
    }
}

// ---------------------------------------------------------------------------
// Render side-outputs of one step (setCell, shaders/compute/operations.glsl:99-170),
// after the step that turned `prev` into `cur`:
//   colour: the cell's material colour; unless EMPTY, rgb minus
//           noise(pos, 3, 2.0, 0.25) * 0.25, clamped to [0, 1] (math.glsl:80-106);
//   light:  the material's emission if its rgb is non-zero, else full light on
//           row 0, else the mix (0.5) of the average and the maximum of the 8
//           neighbours' previous light (neighbours from the pre-step grid; a
//           light obstacle passes only its falloff), in getDiagonalNeighbours
//           order (math.glsl:154-166).
// Float arithmetic follows the GLSL expression order with no contraction; sin
// is this platform's sinf, so the noise matches the GL implementation's only
// to within its sin error.  The noise depends on the position alone: se_noise_fill
// evaluates it once per world into an f32 image that se_render reads.
// ---------------------------------------------------------------------------
__device__ __forceinline__ float se_fract(float x) { return x - floorf(x); }
__device__ __forceinline__ float2 se_old_hash2(float px, float py) {
    const float a = px * 127.1f + py * 311.7f, b = px * 269.5f + py * 183.3f;
    return make_float2(-1.0f + 2.0f * se_fract(sinf(a) * 43758.5453123f),
                       -1.0f + 2.0f * se_fract(sinf(b) * 43758.5453123f));
}

__device__ __forceinline__ float se_noise1(float px, float py) {
    const float K1 = 0.366025404f, K2 = 0.211324865f;
    const float s = (px + py) * K1;
    const float ix = floorf(px + s), iy = floorf(py + s);
    const float t = (ix + iy) * K2;
    const float ax = px - ix + t, ay = py - iy + t;
    const float m = ax < ay ? 0.0f : 1.0f;  // step(a.y, a.x)
    const float ox = m, oy = 1.0f - m;
    const float bx = ax - ox + K2, by = ay - oy + K2;
    const float k2x2 = 2.0f * K2;
    const float cx = ax - 1.0f + k2x2, cy = ay - 1.0f + k2x2;
    const float h0 = fmaxf(0.5f - (ax * ax + ay * ay), 0.0f);
    const float h1 = fmaxf(0.5f - (bx * bx + by * by), 0.0f);
    const float h2 = fmaxf(0.5f - (cx * cx + cy * cy), 0.0f);
    const float2 g0 = se_old_hash2(ix, iy), g1 = se_old_hash2(ix + ox, iy + oy), g2 = se_old_hash2(ix + 1.0f, iy + 1.0f);
    const float n0 = h0 * h0 * h0 * h0 * (ax * g0.x + ay * g0.y);
    const float n1 = h1 * h1 * h1 * h1 * (bx * g1.x + by * g1.y);
    const float n2 = h2 * h2 * h2 * h2 * (cx * g2.x + cy * g2.y);
    return 0.25f + 0.5f * (n0 * 70.0f + n1 * 70.0f + n2 * 70.0f);
}

__device__ __forceinline__ float se_noise(float px, float py) {  // noise(p, 3, 2.0, 0.25)
    float f = 0.0f, qx = px, qy = py;
#pragma unroll
    for (int o = 1; o < 4; o++) {
        f += 1.0f / (float)o * se_noise1(qx * 0.25f, qy * 0.25f);
        qx *= 2.0f;
        qy *= 2.0f;
    }
    return f;
}

// Noise image of the colour output: noise(x, y) for every cell (once per world).
extern "C" __global__ __launch_bounds__(256) void se_noise_fill(float* __restrict__ noise, int W, int H, int pitch) {
    const int x = (int)blockIdx.x * 64 + (int)(threadIdx.x & 63), y = (int)blockIdx.y * 4 + (int)(threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    noise[(size_t)y * pitch + x] = se_noise((float)x, (float)y);
}

// Raw inputs of one row for the light terms: the lane's column and, for lanes
// 0 and 63, the column just outside the wavefront's strip (light, pre-step id).
struct se_lraw {
    float4 L, Le;
    u32 P, Pe;
};

__device__ __forceinline__ se_lraw se_light_load(const u32* __restrict__ prev, const float4* __restrict__ light_in,
                                                 int yy, int x, int lane, int W, int H, int pitch) {
    se_lraw r;
    r.L = r.Le = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    r.P = r.Pe = SE_EMPTY;
    if (yy < 0 || yy >= H) return r;  // uniform across the wavefront
    const size_t row = (size_t)yy * pitch;
    if (x < W) {
        r.L = light_in[row + x];
        r.P = prev[row + x];
    }
    const int xe = lane == 0 ? x - 1 : x + 1;
    if ((lane == 0 || lane == 63) && xe >= 0 && xe < W) {
        r.Le = light_in[row + xe];
        r.Pe = prev[row + xe];
    }
    return r;
}

// One neighbour's light as setCell uses it: (rgb * pass * w, w), pass = 0 for a
// light obstacle in the pre-step grid.  Per column, shared by the 3 cells that
// take it as a neighbour.
__device__ __forceinline__ float4 se_light_term(u32 pid, float4 ld) {
    const float pass = se_light_obstacle(se_decode(pid)) ? 0.0f : 1.0f;
    ld.x *= pass;
    ld.y *= pass;
    ld.z *= pass;
    return make_float4(ld.x * ld.w, ld.y * ld.w, ld.z * ld.w, ld.w);
}

__device__ __forceinline__ float se_dpp_from_left(float v) {  // lane i gets lane i-1 (wave_shr:1)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float se_dpp_from_right(float v) {  // lane i gets lane i+1 (wave_shl:1)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

// Light terms of one row at columns x-1, x, x+1 for the lane's column x: each
// lane computes its own column; x-1 and x+1 come from the neighbouring lanes,
// and lanes 0 and 63 compute the columns just outside the wavefront's strip.
// A row outside the world, and columns outside it, hold zeros (never used).
struct se_lrow {
    float4 l, c, r;
};

__device__ __forceinline__ se_lrow se_light_row(const se_lraw& r, int lane) {
    se_lrow o;
    o.c = se_light_term(r.P, r.L);
    const float4 e = se_light_term(r.Pe, r.Le);
    o.l = make_float4(se_dpp_from_left(o.c.x), se_dpp_from_left(o.c.y), se_dpp_from_left(o.c.z),
                      se_dpp_from_left(o.c.w));
    o.r = make_float4(se_dpp_from_right(o.c.x), se_dpp_from_right(o.c.y), se_dpp_from_right(o.c.z),
                      se_dpp_from_right(o.c.w));
    if (lane == 0) o.l = e;
    if (lane == 63) o.r = e;
    return o;
}

// One neighbour step of setCell's light loop (math.glsl:154-166 order).
__device__ __forceinline__ void se_light_acc(float4 l, bool in, float4& avg, float4& mx, float& max_falloff, int& n) {
    if (!in) return;
    const float falloff = l.w == 0.0f ? max_falloff : l.w;
    avg.x += l.x;
    avg.y += l.y;
    avg.z += l.z;
    avg.w += falloff;
    max_falloff = fmaxf(falloff, max_falloff);
    n++;
    mx.x = fmaxf(mx.x, l.x);
    mx.y = fmaxf(mx.y, l.y);
    mx.z = fmaxf(mx.z, l.z);
    mx.w = fmaxf(mx.w, falloff);
}

// One wavefront per 64-column strip and SE_RENDER_ROWS rows, walking down the
// rows with the light terms of rows y-1, y, y+1 in registers: every input is
// read from HBM about once (the strip's 2 extra rows are mostly L2 hits: the
// strip below is dispatched W/256 workgroups later, on the same XCD when W/256
// is a multiple of 8), and light and colour are written with 16-byte
// non-temporal stores.  Grid: ((W + 255) / 256, ceil(H / SE_RENDER_ROWS))
// workgroups of 4 wavefronts.
#define SE_RENDER_ROWS 8
extern "C" __global__ __launch_bounds__(256) void se_render(const u32* __restrict__ prev, const u32* __restrict__ cur,
                                                            const float4* __restrict__ light_in,
                                                            float4* __restrict__ light_out,
                                                            float4* __restrict__ color_out,
                                                            const float* __restrict__ noise, int W, int H, int pitch) {
    const int lane = (int)(threadIdx.x & 63);
    const int sx = ((int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * 64;
    if (sx >= W) return;
    const int x = sx + lane;
    const int y0 = (int)blockIdx.y * SE_RENDER_ROWS, y1 = min(H, y0 + SE_RENDER_ROWS);
    const bool inl = x - 1 >= 0 && x - 1 < W, inr = x + 1 < W;
    se_lrow up = se_light_row(se_light_load(prev, light_in, y0 - 1, x, lane, W, H, pitch), lane);
    se_lrow mid = se_light_row(se_light_load(prev, light_in, y0, x, lane, W, H, pitch), lane);
    // inputs run one row ahead of their use: row y+2's light and ids, row y+1's
    // id and noise are in flight while row y is computed
    se_lraw nxt = se_light_load(prev, light_in, y0 + 1, x, lane, W, H, pitch);
    u32 cn = SE_EMPTY;
    float nzn = 0.0f;
    if (x < W) {
        cn = cur[(size_t)y0 * pitch + x];
        nzn = noise[(size_t)y0 * pitch + x];
    }
    for (int y = y0; y < y1; y++) {
        const se_lraw nxt2 = se_light_load(prev, light_in, y + 2, x, lane, W, H, pitch);
        const u32 cid = cn;
        const float nz = nzn;
        if (x < W && y + 1 < y1) {
            cn = cur[(size_t)(y + 1) * pitch + x];
            nzn = noise[(size_t)(y + 1) * pitch + x];
        }
        const se_lrow dn = se_light_row(nxt, lane);
        nxt = nxt2;
        if (x < W) {
            const size_t p = (size_t)y * pitch + x;
            const u32 c = se_decode(cid);
            float4 col = se_color(c);
            if (c != SE_EMPTY) {
                const float r = nz * 0.25f;
                col.x = fminf(fmaxf(col.x - r, 0.0f), 1.0f);
                col.y = fminf(fmaxf(col.y - r, 0.0f), 1.0f);
                col.z = fminf(fmaxf(col.z - r, 0.0f), 1.0f);
            }
            se_store_f4(color_out + p, col);
            const float4 em = se_emission(c);
            float4 light;
            if (em.x != 0.0f || em.y != 0.0f || em.z != 0.0f) {
                light = em;
            } else if (y == 0) {
                light = make_float4(1.0f, 1.0f, 1.0f, 0.999999f);
            } else {
                const bool ind = y + 1 < H, inu = y - 1 >= 0;
                float4 avg = make_float4(0.0f, 0.0f, 0.0f, 0.0f), mx = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                float max_falloff = 0.0f;
                int n = 0;
                se_light_acc(dn.c, ind, avg, mx, max_falloff, n);         // ( 0,  1)
                se_light_acc(up.c, inu, avg, mx, max_falloff, n);         // ( 0, -1)
                se_light_acc(dn.l, ind && inl, avg, mx, max_falloff, n);  // (-1,  1)
                se_light_acc(up.l, inu && inl, avg, mx, max_falloff, n);  // (-1, -1)
                se_light_acc(dn.r, ind && inr, avg, mx, max_falloff, n);  // ( 1,  1)
                se_light_acc(up.r, inu && inr, avg, mx, max_falloff, n);  // ( 1, -1)
                se_light_acc(mid.r, inr, avg, mx, max_falloff, n);        // ( 1,  0)
                se_light_acc(mid.l, inl, avg, mx, max_falloff, n);        // (-1,  0)
                if (n == 8) {  // x / 8 == x * 0.125 exactly
                    avg.x *= 0.125f;
                    avg.y *= 0.125f;
                    avg.z *= 0.125f;
                    avg.w *= 0.125f;
                } else if (n > 0) {
                    const float fn = (float)n;
                    avg.x /= fn;
                    avg.y /= fn;
                    avg.z /= fn;
                    avg.w /= fn;
                }
                light = make_float4(avg.x * 0.5f + mx.x * 0.5f, avg.y * 0.5f + mx.y * 0.5f,
                                    avg.z * 0.5f + mx.z * 0.5f, avg.w);
            }
            se_store_f4(light_out + p, light);
        }
        up = mid;
        mid = dn;
    }
}
