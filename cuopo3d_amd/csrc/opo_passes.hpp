// opo_passes.hpp -- the two passes of a z-slice as device functions, shared
// by the single-solve kernels and the batched (many-solve) kernels.
//
//   nl_time_tile  one tile of the fused pass: pointwise RK4 half step(s) of
//                 the three fields, then the along-t line operator
//                 (dispersion/walk-off/absorption, or GDD at a trip's end);
//   xy_plane      Q_x then Q_y on one t-plane of one field.
#pragma once
#include "opo_device.hpp"
#include "opo_kernels.hpp"
#include "opo_rk4.hpp"

namespace opo {

// Blocks are handed to the 8 XCDs round-robin (block b runs on XCD b % 8).
// Remap so that neighbouring tiles, which share cache lines at the tile
// edges, are processed on the same XCD and hit its L2.
__device__ __forceinline__ int xcd_local_tile(int b, int nb) {
    if (nb % 8 != 0) return b;
    return (b % 8) * (nb / 8) + b / 8;
}

// LDS budget of the fused pass: three tiles of NT x TX <= 48 KiB (two or
// more workgroups per CU), at least 4 pixels wide.
template <typename T, int NT> constexpr int tline3_tx() {
    constexpr int per_px = 3 * NT * (int)sizeof(cplx<T>);
    constexpr int tx = (per_px * 4 > 49152) ? 98304 / per_px : 49152 / per_px;
    return tx >= 16 ? 16 : (tx >= 8 ? 8 : (tx >= 4 ? 4 : (tx >= 2 ? 2 : 1)));
}

// LDS layout of one field's tile in the fused pass: time row t (TX
// pixels) starts at t*TX + (t / G)*PAD, G the innermost group length of
// the line operator (N2 of the two-level split, N3 of the three-level one).
// Narrow tiles pad every G rows by TX elements so the rows that the
// innermost phase walks in parallel fall in different banks.
// Rows between two pads of a narrow tile: the innermost group of the line
// operator.  (Padding 2-pixel 2048-sample tiles every 16 rows instead,
// conflict-free in a 64-bank model of the LDS, measured no faster.)
template <typename T, int NT, int TX> constexpr int pad_group() { return inner_len<T, NT>(); }
template <typename T, int NT, int TX> struct TimeTile {
    static constexpr int G = pad_group<T, NT, TX>();
    static constexpr int PAD = (G > 1 && TX < 16) ? TX : 0;
    static constexpr int FT = NT * TX + (NT / G) * PAD;   // elements per field
    __device__ static __forceinline__ int at(int i) { return i + (i / (G * TX)) * PAD; }
};

// Launch variants of the fused pass for one NT (chosen by the autotuner).
// V < 9: one tile per workgroup, V = 3 m + b with tile width m = 0: TXd,
// 1: TXd/2, 2: 2 TXd pixels and b = 0: 256, 1: 512, 2: 384 threads (384
// gives every lane a task in the line-operator phases of 4- and 8-pixel
// tiles).  V >= 9: persistent workgroups (k + 1 per CU), each walking a
// contiguous run of tiles with the next tile prefetched into registers;
// V - 9 = 2 (3 m + b) + k with b = 0: 256, 1: 384, 2: 512 threads.
// FP64 has the one-tile variants only; NT < 64 only variant 0; a variant whose three
// tiles exceed 112 KiB of LDS (persistent: tiles plus staged tables beyond
// 144 KiB) does not exist, nor persistent variants of lines longer than 256
// samples.  (Not offered, measured slower or no faster on every grid: a form
// that keeps whole t-lines in one wave's registers, exchanging through
// wave-private LDS, 15-40 % slower; 768 and 1024 threads per tile, i.e.
// three and four waves per SIMD; persistent workgroups on 512- to
// 2048-sample lines, 23-90 % slower, their prefetch registers spilling
// beside 32- and 64-point register DFTs, also with only the twiddles
// staged so that 4-pixel tiles fit; for 2048-sample lines, 4- and 8-pixel
// tiles with one field in LDS at a time and the other two held in
// registers, 28-240 % slower, the held fields spilling beside the
// three-level line operator; a fused form whose threads load their
// outer-level samples of all three fields straight from HBM into
// registers, take the pointwise step and the outer register DFT before the
// first LDS write and store straight from the inverse outer DFT (two LDS
// round trips and two barriers fewer, bitwise the same results): ~245
// VGPRs, 7-20 % slower on 128- to 2048-sample lines.)
constexpr int kTimeClassic = 9;
constexpr int kTimeVariants = kTimeClassic + 18;
template <typename T, int NT, int V> struct TimeVariant {
    static constexpr int txd = tline3_tx<T, NT>();
    static constexpr bool persistent = V >= kTimeClassic;
    static constexpr int idx = persistent ? V - kTimeClassic : V;
    static constexpr int m = persistent ? idx / 6 : idx / 3;
    static constexpr int b = persistent ? (idx / 2) % 3 : idx % 3;
    static constexpr int per_cu = persistent ? idx % 2 + 1 : 0;
    static constexpr int tx = m == 0 ? txd : (m == 1 ? txd / 2 : txd * 2);
    static constexpr int bs = persistent ? (b == 0 ? 256 : (b == 1 ? 384 : 512)) : (b == 0 ? 256 : (b == 1 ? 512 : 384));
    // one-pixel tiles only where nothing wider fits (FP64 2048-sample lines:
    // there they give the 384- and 512-thread blocks)
    static constexpr int min_tx = ((sizeof(T) == 8 && NT >= 2048) || NT >= 4096) ? 1 : 2;
    static constexpr int data_bytes = 3 * TimeTile<T, NT, (tx > 0 ? tx : 1)>::FT * (int)sizeof(cplx<T>);
    static constexpr int table_bytes = 8 * NT * (int)sizeof(cplx<T>);  // twiddles + 3 tables, hi/lo
    static constexpr bool valid =
        V == 0 || ((sizeof(T) == 4 || !persistent) && NT >= 64 && tx >= min_tx && tx <= 32 && data_bytes <= 114688 &&
                   (!persistent || (NT <= 256 && data_bytes + table_bytes <= 147456)));
};
// Untuned choice: the default tile with 512 threads where it exists.
// (FP64 2048-sample lines: one pixel on 384 threads, 5 % faster than 256.)
constexpr int time_default_variant(bool fp64, int nt) {
    return (!fp64 && nt >= 64) ? 1 : ((fp64 && nt >= 2048) ? 2 : 0);
}

// Variants of the xy pass: V < 3 one plane per workgroup of 256, 512 or
// 1024 threads; for NX >= 64, V >= 3 persistent workgroups (k + 1 per CU)
// walking a run of planes with the next plane prefetched into registers,
// V - 3 = 2 b + k with the same block sizes (at most 16 prefetched
// elements per thread); for 16 <= NX <= 32, V = 3 + b groups of 256 / NX
// planes per workgroup (xy_group) with the same block sizes.  (Not
// offered, measured no faster: all three fields of one time sample per
// workgroup; 512 threads held to 80 registers so that three 64 x 64
// planes share a CU; one line per thread on 64-256 threads for NX <= 32;
// persistent workgroups for NX <= 32; one plane on 384 or 768 threads,
// which balance the 96-point lines' 768 and 1152 tasks per phase.)
constexpr int kXyVariants = 9;
constexpr bool xy_variant_persistent(int nx, int v) { return nx >= 64 && v >= 3; }
constexpr int xy_variant_bs(int nx, int v) {
    const int b = xy_variant_persistent(nx, v) ? (v - 3) / 2 : v % 3;
    return b == 0 ? 256 : (b == 1 ? 512 : 1024);
}
constexpr int xy_variant_per_cu(int nx, int v) { return xy_variant_persistent(nx, v) ? (v - 3) % 2 + 1 : 0; }
constexpr int xy_default_variant(int nx) { return nx >= 96 ? 1 : 0; }

template <typename T> struct alignas(4 * sizeof(T)) cpair {
    cplx<T> a, b;
};

// Values move component-wise: a cpair<double> or cplx<double> copied as a
// struct between memory spaces goes through a private-memory temporary
// (scratch) in the fused-pass kernels.
template <typename T> __device__ __forceinline__ cplx<T> ld_c(const cplx<T> *p) { return cplx<T>{p->x, p->y}; }
template <typename T> __device__ __forceinline__ void st_c(cplx<T> *p, cplx<T> v) {
    p->x = v.x;
    p->y = v.y;
}
// Two consecutive cells of a field in global memory (p 4 sizeof(T)-byte
// aligned: an even cell index), to registers and back from LDS (q).
template <typename T> __device__ __forceinline__ cpair<T> load_pair(const cplx<T> *p) {
    const OPO_GLOBAL cpair<T> *g = (const OPO_GLOBAL cpair<T> *)p;
    return cpair<T>{cplx<T>{g->a.x, g->a.y}, cplx<T>{g->b.x, g->b.y}};
}
template <typename T> __device__ __forceinline__ void store_pair(cplx<T> *p, const cplx<T> *q) {
    OPO_GLOBAL cpair<T> *g = (OPO_GLOBAL cpair<T> *)p;
    const cplx<T> a = ld_c(q), b = ld_c(q + 1);
    g->a.x = a.x;
    g->a.y = a.y;
    g->b.x = b.x;
    g->b.y = b.y;
}

// Pointwise part of the fused pass for two cells: RK4 half step(s), then
// the mirror factors on s and i when the pass closes a round trip.
template <typename T, bool PHASE>
__device__ __forceinline__ void nl_point_pair(const Rk4Args<T> &rk, int t, Tri<T> &A0, Tri<T> &A1) {
    if (rk.nsteps > 0) {
        rk4_half_step<T, PHASE>(rk, 0, A0);
        rk4_half_step<T, PHASE>(rk, 0, A1);
        if (rk.nsteps > 1) {
            rk4_half_step<T, PHASE>(rk, 1, A0);
            rk4_half_step<T, PHASE>(rk, 1, A1);
        }
    }
    if (rk.out_scale) {
        if (rk.eom) {
            const cplx<T> e = rk.eom[2 * t], el = rk.eom[2 * t + 1];
            A0.s = cmul_hl(A0.s, e, el);
            A1.s = cmul_hl(A1.s, e, el);
            A0.i = cmul_hl(A0.i, e, el);
            A1.i = cmul_hl(A1.i, e, el);
        }
        const cplx<T> ms = {rk.ms_re, rk.ms_im}, msl = {rk.ms_lo_re, rk.ms_lo_im};
        const cplx<T> mi = {rk.mi_re, rk.mi_im}, mil = {rk.mi_lo_re, rk.mi_lo_im};
        A0.s = cmul_hl(A0.s, ms, msl);
        A1.s = cmul_hl(A1.s, ms, msl);
        A0.i = cmul_hl(A0.i, mi, mil);
        A1.i = cmul_hl(A1.i, mi, mil);
    }
}

// Pointwise part of the fused pass for one cell at time sample t (the same
// operations as nl_point_pair applies to each of its two cells).
template <typename T, bool PHASE>
__device__ __forceinline__ void nl_point(const Rk4Args<T> &rk, int t, Tri<T> &A) {
    if (rk.nsteps > 0) rk4_half_step<T, PHASE>(rk, 0, A);
    if (rk.nsteps > 1) rk4_half_step<T, PHASE>(rk, 1, A);
    if (rk.out_scale) {
        if (rk.eom) {
            const cplx<T> e = rk.eom[2 * t], el = rk.eom[2 * t + 1];
            A.s = cmul_hl(A.s, e, el);
            A.i = cmul_hl(A.i, e, el);
        }
        A.s = cmul_hl(A.s, cplx<T>{rk.ms_re, rk.ms_im}, cplx<T>{rk.ms_lo_re, rk.ms_lo_im});
        A.i = cmul_hl(A.i, cplx<T>{rk.mi_re, rk.mi_im}, cplx<T>{rk.mi_lo_re, rk.mi_lo_im});
    }
}

// Whether a one-tile workgroup of the fused pass copies the NT-point
// twiddles (2 NT elements, hi/lo) into LDS: the line operator then reads
// them from LDS instead of every lane fetching its (hi, lo) pairs through
// the vector memory path, which for 32- and 16-point register DFTs is the
// larger part of a phase (128 x 128 x 1024 t pass 462 -> 432 us; cavity
// 64 x 64 x 256 +1.2 %, 32 x 32 x 128 +2.5 %).  Only for lines of 128 to
// 1024 samples (two-level, and the three-level 1024-sample lines, measured
// with the copy) and where the copy fits beside the tile (64-sample lines
// measured no faster, the three-level 2048-sample lines 4 % slower; copying
// the line tables as well measured 0.1-1.3 % slower).
template <typename T, int NT, int TX> constexpr bool time_stage_twiddles() {
    return NT >= 128 && NT <= 1024 &&
           (3 * TimeTile<T, NT, TX>::FT + 2 * NT) * (int)sizeof(cplx<T>) <= 150 * 1024;
}

// One tile of the fused pass: TX consecutive pixels of row y, all NT time
// samples, all three fields.  lds holds 3 * NT * TX elements; twl holds the
// 2 NT staged twiddles where time_stage_twiddles() (else unused).
template <typename T, int NT, int TX, int BS, bool PHASE>
__device__ __forceinline__ void nl_time_tile(const TLine3Args<T> &a, int nx, int ny, int tile,
                                             const cplx<T> *__restrict__ tw, cplx<T> *lds, cplx<T> *twl) {
    using L = TimeTile<T, NT, TX>;
    constexpr int TILE = NT * TX, FT = L::FT;
    constexpr bool STW = time_stage_twiddles<T, NT, TX>();
    constexpr int TWP = STW ? (2 * NT + BS - 1) / BS : 1;
    const int tid = threadIdx.x;
    // twiddle copy: loads issued now, stored to LDS after the pointwise step
    cplx<T> twr[TWP];
    if constexpr (STW) {
#pragma unroll
        for (int j = 0; j < TWP; ++j)
            if (tid + j * BS < 2 * NT) twr[j] = gld_c(tw + tid + j * BS);
    }
    const int tiles_x = nx / TX;
    const int y = tile / tiles_x, x0 = (tile % tiles_x) * TX;
    const size_t plane = (size_t)nx * ny;
    const size_t base = (size_t)y * nx + x0;
    const Rk4Args<T> &rk = a.rk;
    // pointwise step in registers, two pixels per thread per iteration
    // (versions that issued the loads of one or of four iterations ahead of
    // the arithmetic were measured no faster: more registers, no latency
    // gained)
    if constexpr (TX % 2 == 0) {
        for (int i = 2 * tid; i < TILE; i += 2 * BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            cpair<T> p = load_pair(a.psrc + gi);
            cpair<T> s = load_pair(a.f[1] + gi);
            cpair<T> q = load_pair(a.f[2] + gi);
            Tri<T> A0 = {p.a, s.a, q.a}, A1 = {p.b, s.b, q.b};
            nl_point_pair<T, PHASE>(rk, t, A0, A1);
            const int li = L::at(i);
            lds[li] = A0.p;
            lds[li + 1] = A1.p;
            lds[FT + li] = A0.s;
            lds[FT + li + 1] = A1.s;
            lds[2 * FT + li] = A0.i;
            lds[2 * FT + li + 1] = A1.i;
        }
    } else {
        for (int i = tid; i < TILE; i += BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            Tri<T> A = {gld_c(a.psrc + gi), gld_c(a.f[1] + gi), gld_c(a.f[2] + gi)};
            nl_point<T, PHASE>(rk, t, A);
            const int li = L::at(i);
            lds[li] = A.p;
            lds[FT + li] = A.s;
            lds[2 * FT + li] = A.i;
        }
    }
    if constexpr (STW) {
#pragma unroll
        for (int j = 0; j < TWP; ++j)
            if (tid + j * BS < 2 * NT) st_c(twl + tid + j * BS, twr[j]);
    }
    __syncthreads();
    const LineSet<T> ls{lds + a.first_field * FT, FT, a.tab0, a.tab_stride, 3 - a.first_field};
    // the line tables are in global memory (measured: the batched 512-sample
    // t pass 0.5 % faster than with generic reads; the xy planes were slower)
    line_op_any<T, NT, +1, 1, TX, L::PAD, false, L::G, true>(ls, TX, STW ? twl : tw, tid, BS);
    if constexpr (TX % 2 == 0) {
        for (int i = 2 * tid; i < TILE; i += 2 * BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            const int li = L::at(i);
            store_pair(a.f[0] + gi, lds + li);
            store_pair(a.f[1] + gi, lds + FT + li);
            store_pair(a.f[2] + gi, lds + 2 * FT + li);
        }
    } else {
        for (int i = tid; i < TILE; i += BS) {
            const int t = i / TX, xl = i % TX;
            const size_t gi = (size_t)t * plane + base + xl;
            for (int f = 0; f < 3; ++f) gst_c(a.f[f] + gi, ld_c(lds + f * FT + L::at(i)));
        }
    }
}



// Contiguous run [u0, u1) of the work units of this workgroup.
__device__ __forceinline__ void unit_run(int units, int &u0, int &u1) {
    const long long G = gridDim.x, b = blockIdx.x;
    u0 = (int)(b * units / G);
    u1 = (int)((b + 1) * units / G);
}

// Persistent form of the fused pass: the workgroup walks a run of tiles.
// The raw fields of tile u+1 are loaded into registers as soon as tile u
// is in LDS, so their HBM latency hides behind the line operator of tile
// u; twiddles and tables are staged in LDS (stab, 8 NT elements) whenever
// the tile belongs to another solve than the staged tables, so the
// prefetch is the only global load in flight inside the loop.
//
// Src supplies the work units: units(), skip(u), key(u) (identity of the
// tables a unit needs), args(u) (pointers, step weights, first field),
// nx(u), ny(u), tile(u), and stage(u, stab, tid, nthr) (copy twiddles to
// stab[0, 2NT) and the line tables of fields first..2 after them).
template <typename T, int NT, int TX, int BS, typename Src>
__device__ __forceinline__ void nl_time_run(const Src &src, cplx<T> *lds, cplx<T> *stab) {
    static_assert(TX % 2 == 0, "pairs of pixels");
    using L = TimeTile<T, NT, TX>;
    constexpr int TILE = NT * TX, FT = L::FT;
    constexpr int PT = (TILE + 2 * BS - 1) / (2 * BS);
    constexpr bool EVEN = TILE % (2 * BS) == 0;
    const int tid = threadIdx.x;
    int u0, u1;
    unit_run(src.units(), u0, u1);
    cpair<T> rp[PT], rs[PT], ri[PT];
    auto offset = [&](int u, int i) {
        const int nx = src.nx(u), tiles_x = nx / TX, tile = src.tile(u);
        const size_t base = (size_t)(tile / tiles_x) * nx + (tile % tiles_x) * TX;
        return (size_t)(i / TX) * ((size_t)nx * src.ny(u)) + base + i % TX;
    };
    auto load = [&](int u) {
        if (src.skip(u)) return;
        const TLine3Args<T> a = src.args(u);
#pragma unroll
        for (int j = 0; j < PT; ++j) {
            const int i = 2 * tid + j * 2 * BS;
            if (EVEN || i < TILE) {
                const size_t gi = offset(u, i);
                rp[j] = load_pair(a.psrc + gi);
                rs[j] = load_pair(a.f[1] + gi);
                ri[j] = load_pair(a.f[2] + gi);
            }
        }
    };
    int staged = -1;
    if (u0 < u1) load(u0);
    for (int u = u0; u < u1; ++u) {
        if (src.skip(u)) {
            if (u + 1 < u1) load(u + 1);
            continue;
        }
        const TLine3Args<T> a = src.args(u);
        if (src.key(u) != staged) {   // the barrier below publishes the tables
            src.stage(u, stab, tid, BS);
            staged = src.key(u);
        }
#pragma unroll
        for (int j = 0; j < PT; ++j) {
            const int i = 2 * tid + j * 2 * BS;
            if (EVEN || i < TILE) {
                Tri<T> A0 = {rp[j].a, rs[j].a, ri[j].a}, A1 = {rp[j].b, rs[j].b, ri[j].b};
                nl_point_pair<T, false>(a.rk, i / TX, A0, A1);
                const int li = L::at(i);
                lds[li] = A0.p;
                lds[li + 1] = A1.p;
                lds[FT + li] = A0.s;
                lds[FT + li + 1] = A1.s;
                lds[2 * FT + li] = A0.i;
                lds[2 * FT + li + 1] = A1.i;
            }
        }
        __syncthreads();
        if (u + 1 < u1) load(u + 1);
        const LineSet<T> ls{lds + a.first_field * FT, FT, stab + 2 * NT, 2 * NT, 3 - a.first_field};
        line_op_any<T, NT, +1, 1, TX, L::PAD, false, L::G>(ls, TX, stab, tid, BS);
        for (int i = 2 * tid; i < TILE; i += 2 * BS) {
            const size_t gi = offset(u, i);
            const int li = L::at(i);
            store_pair(a.f[0] + gi, lds + li);
            store_pair(a.f[1] + gi, lds + FT + li);
            store_pair(a.f[2] + gi, lds + 2 * FT + li);
        }
        __syncthreads();
    }
}

// Copy twiddles (2NT elements) and the line tables of fields first..2
// (each 2NT elements, tab_stride apart) into stab.
template <typename T, int NT>
__device__ __forceinline__ void stage_time_tables(const cplx<T> *__restrict__ tw, const TLine3Args<T> &a,
                                                  cplx<T> *stab, int tid, int nthr) {
    // the four arrays' loads of an index go out together (one memory
    // latency per index instead of one per array)
    const int nf = 3 - a.first_field;
    for (int i = tid; i < 2 * NT; i += nthr) {
        cplx<T> v[4];
        v[0] = gld_c(tw + i);
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (f < nf) v[1 + f] = gld_c(a.tab0 + (size_t)f * a.tab_stride + i);
        stab[i] = v[0];
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (f < nf) stab[2 * NT * (1 + f) + i] = v[1 + f];
    }
}

// Work units of the fused pass of one solve: its tiles.
template <typename T, int NT, int TX> struct TimeSolo {
    const TLine3Args<T> &a;
    int nx_, ny_;
    const cplx<T> *tw;
    __device__ int units() const { return (nx_ / TX) * ny_; }
    __device__ bool skip(int) const { return false; }
    __device__ int key(int) const { return 0; }
    __device__ const TLine3Args<T> &args(int) const { return a; }
    __device__ int nx(int) const { return nx_; }
    __device__ int ny(int) const { return ny_; }
    __device__ int tile(int u) const { return u; }
    __device__ void stage(int, cplx<T> *stab, int tid, int nthr) const {
        stage_time_tables<T, NT>(tw, a, stab, tid, nthr);
    }
};

// Whether the xy line operators of an NX plane on BS threads read their
// constants wave-uniformly (through the scalar cache).
template <int NX, int BS> constexpr bool xy_uniform() { return NX % 64 == 0 && BS % 64 == 0; }

// Q_x then Q_y on P consecutive NX x NX planes starting at A (x fastest,
// planes NX*NX apart).  lds holds P*NX*(NX+1)
// elements: rows padded by one element so row and column walks spread over
// the LDS banks.  The planes share the tables (one field); each line goes
// through exactly the operations it would alone.
template <typename T, int NX, int BS, int P = 1>
__device__ __forceinline__ void xy_plane(cplx<T> *A, const cplx<T> *__restrict__ tabx,
                                         const cplx<T> *__restrict__ taby,
                                         const cplx<T> *__restrict__ tw, cplx<T> *lds) {
    constexpr int NXP = NX + 1, NE = P * NX * NX;
    const int tid = threadIdx.x;
    chunked_copy<T, NE, BS>(tid, [&](int i) { return gld_c(A + i); },
                            [&](int i, cplx<T> v) { lds[(i / NX) * NXP + (i % NX)] = v; });
    __syncthreads();
    constexpr bool UNI = xy_uniform<NX, BS>();
    // Q_x: lines are rows (pitch NXP), elements contiguous.
    const LineSet<T> lx{lds, NX * NXP, tabx, 0, P};
    line_op_any<T, NX, -1, NXP, 1, 0, UNI>(lx, NX, tw, tid, BS);
    // Q_y: lines are columns, elements one row apart.
    const LineSet<T> ly{lds, NX * NXP, taby, 0, P};
    line_op_any<T, NX, -1, 1, NXP, 0, UNI>(ly, NX, tw, tid, BS);
    chunked_copy<T, NE, BS>(tid, [&](int i) { return lds[(i / NX) * NXP + (i % NX)]; },
                            [&](int i, cplx<T> v) { gst_c(A + i, v); });
}

// Persistent form of the xy pass: the workgroup walks a run of planes and
// loads plane u+1 into registers while plane u is transformed.  Where the
// constants are read wave-uniformly they come from global memory through
// the scalar cache; otherwise stab holds 14 NX staged elements: twiddles,
// then the x tables of three fields, then their y tables.  Src supplies
// units(), skip(u), key(u), plane(u) (the plane's first element), slot(u)
// (which staged table pair it uses), stage(u, stab, tid, nthr) and
// tables(u, tw, tx, ty) (the global twiddles and tables of unit u).
template <typename T, int NX, int BS, typename Src>
__device__ __forceinline__ void xy_run(const Src &src, cplx<T> *lds, cplx<T> *stab) {
    constexpr bool UNI = xy_uniform<NX, BS>();
    constexpr int NXP = NX + 1, NN = NX * NX;
    constexpr int PT = (NN + BS - 1) / BS;
    constexpr bool EVEN = NN % BS == 0;
    const int tid = threadIdx.x;
    int u0, u1;
    unit_run(src.units(), u0, u1);
    cplx<T> r[PT];
    auto load = [&](int u) {
        if (src.skip(u)) return;
        const cplx<T> *A = src.plane(u);
#pragma unroll
        for (int j = 0; j < PT; ++j) {
            const int i = tid + j * BS;
            if (EVEN || i < NN) r[j] = gld_c(A + i);
        }
    };
    int staged = -1;
    if (u0 < u1) load(u0);
    for (int u = u0; u < u1; ++u) {
        if (src.skip(u)) {
            if (u + 1 < u1) load(u + 1);
            continue;
        }
        if (!UNI && src.key(u) != staged) {
            src.stage(u, stab, tid, BS);
            staged = src.key(u);
        }
#pragma unroll
        for (int j = 0; j < PT; ++j) {
            const int i = tid + j * BS;
            if (EVEN || i < NN) lds[(i / NX) * NXP + (i % NX)] = r[j];
        }
        __syncthreads();
        if (u + 1 < u1) load(u + 1);
        if constexpr (UNI) {
            const cplx<T> *tw, *tx, *ty;
            src.tables(u, tw, tx, ty);
            line_op<T, NX, -1, NXP, 1, true>(lds, NX, tw, tx, tid, BS);
            line_op<T, NX, -1, 1, NXP, true>(lds, NX, tw, ty, tid, BS);
        } else {
            const int k = src.slot(u);
            line_op<T, NX, -1, NXP, 1>(lds, NX, stab, stab + 2 * NX * (1 + k), tid, BS);
            line_op<T, NX, -1, 1, NXP>(lds, NX, stab, stab + 2 * NX * (4 + k), tid, BS);
        }
        cplx<T> *A = src.plane(u);
        for (int i = tid; i < NN; i += BS) gst_c(A + i, lds[(i / NX) * NXP + (i % NX)]);
        __syncthreads();
    }
}

// Copy twiddles and nf pairs of (x, y) tables into stab.
template <typename T, int NX>
__device__ __forceinline__ void stage_xy_tables(const cplx<T> *__restrict__ tw, const cplx<T> *const *tx,
                                                const cplx<T> *const *ty, int nf, cplx<T> *stab, int tid,
                                                int nthr) {
    // all seven arrays' loads of an index go out together
    for (int i = tid; i < 2 * NX; i += nthr) {
        cplx<T> w = gld_c(tw + i), vx[3], vy[3];
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (f < nf) {
                vx[f] = gld_c(tx[f] + i);
                vy[f] = gld_c(ty[f] + i);
            }
        stab[i] = w;
#pragma unroll
        for (int f = 0; f < 3; ++f)
            if (f < nf) {
                stab[2 * NX * (1 + f) + i] = vx[f];
                stab[2 * NX * (4 + f) + i] = vy[f];
            }
    }
}

// Work units of the xy pass of one solve: plane t of field f, u = f nt + t.
template <typename T, int NX> struct XySolo {
    const FieldOps<T> &ops;
    int nt, nf;
    const cplx<T> *tw;
    __device__ int units() const { return nt * nf; }
    __device__ bool skip(int) const { return false; }
    __device__ int key(int) const { return 0; }
    __device__ int slot(int u) const { return u / nt; }
    __device__ cplx<T> *plane(int u) const {
        const int f = u / nt;
        cplx<T> *base = f == 0 ? ops.f[0] : (f == 1 ? ops.f[1] : ops.f[2]);
        return base + (size_t)(u % nt) * NX * NX;
    }
    __device__ void stage(int, cplx<T> *stab, int tid, int nthr) const {
        stage_xy_tables<T, NX>(tw, ops.tab, ops.taby, nf, stab, tid, nthr);
    }
    __device__ void tables(int u, const cplx<T> *&tw_, const cplx<T> *&tx, const cplx<T> *&ty) const {
        const int f = u / nt;
        tw_ = tw;
        tx = f == 0 ? ops.tab[0] : (f == 1 ? ops.tab[1] : ops.tab[2]);
        ty = f == 0 ? ops.taby[0] : (f == 1 ? ops.taby[1] : ops.taby[2]);
    }
};

}  // namespace opo
