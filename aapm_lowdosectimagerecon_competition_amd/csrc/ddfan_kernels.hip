// ddfan_kernels.hip -- gfx950 kernels of the DD fan-beam path.
//
//   k_sat_rows   summed-area tables along one axis (volume rows along x or y
//                with the mask fused, or sinogram rows along the detector)
//                (MultiSlicesProjGPU.cu:26-97,202-209; MultiSlicesBackGPU.cu:12-66)
//   k_proj       branchless DD projector (MultiSlicesProjGPU.cu:100-200) with an
//                optional OS-SART ratio epilogue (OSSART_AAPM.m:81-88)
//   k_back       voxel-driven DD backprojector (MultiSlicesBackGPU.cu:115-184)
//                with optional OS-SART normalise/update and FISTA epilogues
//                (OSSART_AAPM.m:91-97,111-115)
//   k_rebin      helical-to-fan gather (HelicalToFanFunc_mex.cpp:41-139)
//
// No texture units.  A summed-area-table entry k holds the prefix S[k] (sum of
// samples 0..k-1) next to the sample value v[k] (v[N] = 0), so the
// reference's clamp-to-edge linear fetch T(c) = S[i] + t*(S[i+1] - S[i]) is a
// single fma on one paired load: T = S[i0] + t*v[i0] with i0 = clamp(i, 0, N)
// and t = 0 where the clamp bites.  Tables are stored [chunk][row][zt][k]:
// the ZT slices of a chunk share one view angle, so the per-ray index
// arithmetic runs once per chunk and lanes read consecutive k.  ZT = 1 is the
// general per-slice-angle path.  Geometry that decides an integer goes through
// ddfan_index.h and is bit-identical to the host restatement.
#include <algorithm>

#include "ddfan_index.h"
#include "ddfan_internal.h"

namespace ddf {

// Table entry of each precision variant: (S, v), (S_hi, S_lo, v, 0), (S, v).
template <int MODE>
struct Sat;
template <>
struct Sat<SAT_F32> {
    using E = float2;
};
template <>
struct Sat<SAT_TWOFLOAT> {
    using E = float4;
};
template <>
struct Sat<SAT_F64> {
    using E = double2;
};

// gfx950 hands consecutive workgroups to its 8 XCDs round-robin.  Blocks that
// share an XCD (and its L2) get a contiguous range of linear work items, so
// the concurrently resident blocks of one XCD sweep the same slice chunk.
__device__ __forceinline__ int xcd_linear(int b, int nblocks) {
    const int per = nblocks >> 3;  // grids are padded to a multiple of 8
    return (b & 7) * per + (b >> 3);
}
static inline int xcd_grid(int nwork) { return ((nwork + 7) / 8) * 8; }

// Value of the same register in lane - 1 (lane 0 reads 0): DPP wave_shr:1.
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ double lane_prev(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)(b & 0xffffffffu), 0x138, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), 0x138, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Fetch position of a texture coordinate c on a table with entries 0..N:
// i0 = clamp(floor(c - 0.5), 0, N) (also as a 32-bit byte offset of entry
// type E) and the weight t = clamp(c - 0.5 - i0, 0, 1).  This is the
// clamp-to-edge fetch: below the first texel t clamps to 0 and the entry
// reads S[0]; at or above the last, v[N] = 0 makes t irrelevant; in between
// t = (c - 0.5) - floor(c - 0.5) exactly; a NaN coordinate reads entry 0
// with t = 0.  Rows are addressed as (uniform row base) + (32-bit lane
// offset) so loads use the scalar-base addressing form.  The clamp to [0, N]
// is one v_med3_f32, which gives the bits of fminf(fmaxf(x, 0), N) for every
// float x, NaN and signed zeros included (checked on the GPU over all 2^32
// patterns).
struct Fetch {
    unsigned off;
    int i0;
    float t;
};
template <class E>
__device__ __forceinline__ Fetch fetch_pos(float c, int N) {
    const float cm = c - 0.5f;
    const float fc = __builtin_amdgcn_fmed3f(floorf(cm), 0.0f, (float)N);
    Fetch f;
    f.i0 = (int)fc;
    f.off = (unsigned)f.i0 * (unsigned)sizeof(E);
    f.t = fminf(fmaxf(cm - fc, 0.0f), 1.0f);
    return f;
}
template <class E>
__device__ __forceinline__ E entry_at(const E* __restrict__ row, unsigned off) {
    return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(row) + off);
}

// ------------------------------------------------------------------ SAT
template <int MODE>
struct Scan;
template <>
struct Scan<SAT_F32> {
    using L = float;
    float a = 0.0f;
    __device__ __forceinline__ void add(float v) { a = a + v; }
    __device__ __forceinline__ L get() const { return a; }
    __device__ static __forceinline__ L zero() { return 0.0f; }
    __device__ static __forceinline__ L plus(L x, L y) { return x + y; }
    __device__ static __forceinline__ float2 entry(L s, float v) { return make_float2(s, v); }
};
template <>
struct Scan<SAT_TWOFLOAT> {
    using L = float2;  // unnormalised (hi, lo); value = hi + lo
    float h = 0.0f, l = 0.0f;
    __device__ static __forceinline__ float2 two_sum(float a, float b) {
        const float s = a + b;
        const float bb = s - a;
        return make_float2(s, (a - (s - bb)) + (b - bb));
    }
    __device__ __forceinline__ void add(float v) {
        const float2 t = two_sum(h, v);
        h = t.x;
        l = l + t.y;
    }
    __device__ __forceinline__ L get() const { return make_float2(h, l); }
    __device__ static __forceinline__ L zero() { return make_float2(0.0f, 0.0f); }
    __device__ static __forceinline__ L plus(L x, L y) {
        const float2 t = two_sum(x.x, y.x);
        return make_float2(t.x, (x.y + y.y) + t.y);
    }
    __device__ static __forceinline__ float4 entry(L s, float v) {
        const float H = s.x + s.y;  // FastTwoSum normalisation
        return make_float4(H, s.y - (H - s.x), v, 0.0f);
    }
};
template <>
struct Scan<SAT_F64> {
    using L = double;
    double a = 0.0;
    __device__ __forceinline__ void add(float v) { a = a + (double)v; }
    __device__ __forceinline__ L get() const { return a; }
    __device__ static __forceinline__ L zero() { return 0.0; }
    __device__ static __forceinline__ L plus(L x, L y) { return x + y; }
    __device__ static __forceinline__ double2 entry(L s, float v) {
        return make_double2(s, (double)v);
    }
};

// Summed-area tables of one axis.  Sample (row, k, z) sits at
// in[(row*rs + k*ks)*SLN + z], optionally times mask[row*rs + k*ks] (its byte
// value, MultiSlicesProjGPU.cu:202-209).  One block scans one row for G
// consecutive slices: 256/G threads per slice own contiguous k segments
// (loads coalesce over the slices, contiguous in memory), the segment totals
// of a slice are scanned across lanes (lanes G apart) and across the block's
// four waves, and finished rows leave through LDS as contiguous k runs of
// paired entries.  A launch holds several tables (both axes of a
// volume, or the tables of several reconstructions), block ranges one after
// another.
struct SatDev {
    const float* in;
    const uint8_t* mask;
    void* out;
    int nrows, Nk, rs, ks, SLN, ZT, nzc, nzg, blk0;
    int tiled;  // 1: volume table, built by k_sat_tiles when it has >= 16 slices
};
struct SatLaunch {
    int njobs, nwork;
    int nt;  // k_sat_tiles: entries leave as non-temporal stores
    SatDev j[2 * kBatch];
};

// Table entries leave the row builder as non-temporal stores when its launch
// writes at least kSatNtBytes (the size of MI355X's last-level cache): such
// tables are written once, too large to stay cached for the next kernel.
// Measured on MI355X, 512x512x64, 2304 views, two-float: sinogram table
// 0.536 -> 0.502 ms.  Smaller tables keep ordinary stores, which leave
// them cached for the kernel that reads them: 16 reconstructions of 256x256 x
// 8-16 slices x 720 views one after another 0.384 s, 0.451 s with
// non-temporal stores; OS-SART subset tables (72 MB) the same either way.
// The tiled builder takes them from 1 GiB of entries per launch: both volume
// tables of 512x512x200 in one-slice chunks (1.68 GB, two-float) 0.535 ->
// 0.497 ms, and the projector reading them 2.26 -> 2.23 ms; at 512x512x64
// (0.55 GB) 0.141 -> 0.166 ms (0.155 -> 0.215 ms at 32 slices per block), so
// smaller launches keep ordinary stores.
constexpr size_t kSatNtBytes = size_t(256) << 20;
constexpr size_t kSatTileNtBytes = size_t(1) << 30;
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_nt(float4* p, float4 v) {
    __builtin_nontemporal_store(nt_f4{v.x, v.y, v.z, v.w}, reinterpret_cast<nt_f4*>(p));
}
__device__ __forceinline__ void store_nt(float2* p, float2 v) {
    __builtin_nontemporal_store(nt_f2{v.x, v.y}, reinterpret_cast<nt_f2*>(p));
}
__device__ __forceinline__ void store_nt(double2* p, double2 v) {
    __builtin_nontemporal_store(nt_d2{v.x, v.y}, reinterpret_cast<nt_d2*>(p));
}

// Value of lane - d of the same wave (callers use it for lane >= d only).
__device__ __forceinline__ float lane_up(float v, int d) { return __shfl_up(v, d); }
__device__ __forceinline__ double lane_up(double v, int d) { return __shfl_up(v, d); }
__device__ __forceinline__ float2 lane_up(float2 v, int d) {
    return make_float2(__shfl_up(v.x, d), __shfl_up(v.y, d));
}

// Longest scan segment a thread holds in registers: the rows sat_group admits
// (G * K * (prefix + sample bytes) + 256 prefixes within 64 KiB) split over
// 256 / G segments give at most 32 samples (f32 tables, G = 1, K = 8064).
constexpr int kSatSeg = 32;

template <int MODE, int G, bool NT>
__global__ __launch_bounds__(256) void k_sat_rows(const SatLaunch L) {
    using S = Scan<MODE>;
    using Lt = typename S::L;
    using E = typename Sat<MODE>::E;
    constexpr int NSEG = 256 / G;
    extern __shared__ __align__(16) unsigned char smem[];
    // The blocks of one row (its slice groups, which share the input's cache
    // lines) run back to back on one XCD.
    const int bid = xcd_linear(blockIdx.x, gridDim.x);
    if (bid >= L.nwork) return;  // whole block
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (bid >= L.j[t].blk0) ji = t;
    const SatDev& jb = L.j[ji];
    const int local = bid - jb.blk0;
    const int K = jb.Nk + 1;
    const int row = local / jb.nzg;
    const int zg = local - row * jb.nzg;
    const int tz = threadIdx.x % G, seg = threadIdx.x / G;
    const int z = zg * G + tz;
    const bool real = z < jb.SLN;
    Lt* pre = reinterpret_cast<Lt*>(smem);  // [G][K] exclusive prefixes
    Lt* wtot = pre + (size_t)G * K;         // [4 waves][G] wave totals (room for 256)
    float* val = reinterpret_cast<float*>(wtot + (size_t)G * NSEG);  // [G][K] samples
    const int Lseg = (jb.Nk + NSEG - 1) / NSEG;  // <= kSatSeg (launch_sat_group)
    const int k0 = min(seg * Lseg, jb.Nk), k1 = min(k0 + Lseg, jb.Nk);
    // the segment's samples are loaded first, all in flight together, then scanned
    float sv[kSatSeg];
#pragma unroll
    for (int q = 0; q < kSatSeg; ++q) {
        const int k = k0 + q;
        sv[q] = 0.0f;
        if (real && k < k1) {
            const size_t p = (size_t)row * jb.rs + (size_t)k * jb.ks;
            sv[q] = jb.in[p * jb.SLN + z];
            if (jb.mask) sv[q] = sv[q] * (float)jb.mask[p];
        }
    }
    S acc;
#pragma unroll
    for (int q = 0; q < kSatSeg; ++q) {
        const int k = k0 + q;
        if (k < k1) {
            pre[(size_t)tz * K + k] = acc.get();
            val[(size_t)tz * K + k] = sv[q];
            acc.add(sv[q]);
        }
    }
    if (seg == NSEG - 1) val[(size_t)tz * K + jb.Nk] = 0.0f;
    // Exclusive prefix of the segment totals of slice tz: a scan over the
    // wave's lanes of that slice (lane = G * (segment within the wave) + tz),
    // then the totals of the waves before this one.  Every lane takes part.
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Lt incl = acc.get();
#pragma unroll
    for (int d = G; d < 64; d *= 2) {
        const Lt o = lane_up(incl, d);
        if (lane >= d) incl = S::plus(o, incl);
    }
    Lt excl = lane_up(incl, G);
    if (lane < G) excl = S::zero();
    if (lane >= 64 - G) wtot[wave * G + tz] = incl;
    __syncthreads();
    Lt off = S::zero();
    for (int w = 0; w < wave; ++w) off = S::plus(off, wtot[w * G + tz]);
    off = S::plus(off, excl);
    for (int k = k0; k < k1; ++k) pre[(size_t)tz * K + k] = S::plus(pre[(size_t)tz * K + k], off);
    if (seg == NSEG - 1) pre[(size_t)tz * K + jb.Nk] = S::plus(off, acc.get());
    __syncthreads();
    // Rows out, one slice at a time: consecutive threads write consecutive k.
    const int SLNp = jb.nzc * jb.ZT;
    E* base = reinterpret_cast<E*>(jb.out);
#pragma unroll
    for (int q = 0; q < G; ++q) {
        const int zz = zg * G + q;
        if (zz >= SLNp) break;
        const int zc = zz / jb.ZT, zt = zz - zc * jb.ZT;
        E* dst = base + (((size_t)zc * jb.nrows + row) * jb.ZT + zt) * K;
        const Lt* pq = pre + (size_t)q * K;
        const float* vq = val + (size_t)q * K;
        for (int k = threadIdx.x; k < K; k += 256) {
            if constexpr (NT) store_nt(dst + k, S::entry(pq[k], vq[k]));
            else dst[k] = S::entry(pq[k], vq[k]);
        }
    }
}

// Summed-area tables of many slices at once.  A block owns one table row for
// G consecutive slices and walks the row in chunks of KC = (256 / G) * 4
// samples: the chunk's samples are loaded with the G slices of one sample
// position contiguous (4G bytes per position, whole cache lines at G = 32),
// the loads of the next chunk in flight while the current one is scanned
// (its mask bytes are applied when it reaches LDS), transposed through LDS, scanned by 256 / G parts of 4 samples per slice
// (each part sums its 4 samples in order, then adds the slice's running total
// and the totals of the parts before it, in order), and the chunk's entries
// leave through LDS as contiguous runs of each slice's row.  The sums are one
// sequential chain of 4-sample part totals per slice whatever G is (a chunk's
// carry is the chain up to its end), and the builder is picked from the
// table's own slice count, so a table's entries do not depend on G or on
// what shares its launch.  The chunk's samples (tin) are dead once scanned,
// before its entries (tout) are written, so the two share LDS: 19.2 KB per
// block at G = 32, 18.8 KB at G = 16 (two-float), eight blocks per CU, with
// the register budget of eight waves per SIMD.
// V4 (volumes whose slice count is a multiple of 4, 16-byte aligned): each
// thread loads four slices of one position with one 16-byte load, the whole
// chunk in one pass instead of four rounds of 4-byte loads; the same samples
// reach LDS, so the tables are the same.
template <int MODE, int G, bool V4>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_sat_tiles(const SatLaunch L) {
    using S = Scan<MODE>;
    using Lt = typename S::L;
    using E = typename Sat<MODE>::E;
    constexpr int P = 256 / G;  // parts per slice
    constexpr int LP = 4;       // samples per part
    constexpr int KC = P * LP;  // samples per chunk
    constexpr size_t kIn = sizeof(float) * KC * (G + 1), kOut = sizeof(E) * G * (KC + 1);
    __shared__ __align__(16) unsigned char tio[kIn > kOut ? kIn : kOut];
    float(*const tin)[G + 1] = reinterpret_cast<float(*)[G + 1]>(tio);
    E(*const tout)[KC + 1] = reinterpret_cast<E(*)[KC + 1]>(tio);
    __shared__ Lt tot[P][G];
    __shared__ Lt carry[G];
    const int bid = xcd_linear(blockIdx.x, gridDim.x);
    if (bid >= L.nwork) return;  // whole block
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (bid >= L.j[t].blk0) ji = t;
    const SatDev& jb = L.j[ji];
    const int local = bid - jb.blk0;
    const int K = jb.Nk + 1;
    const int row = local / jb.nzg;
    const int z0 = (local - row * jb.nzg) * G;
    const int SLNp = jb.nzc * jb.ZT;
    const int tid = threadIdx.x;
    const int lk = tid / G, lz = tid - lk * G;  // loader: position lk of a round, slice lz
    const int part = tid / G, sz = tid - part * G;  // scanner: part of slice sz
    const bool zin = z0 + lz < jb.SLN;
    if (tid < G) carry[tid] = S::zero();
    E* const base = reinterpret_cast<E*>(jb.out);
    float nv[LP];
    uint8_t nm[LP];
    auto fetch = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < LP; ++r) {
            const int k = k0 + r * P + lk;
            nv[r] = 0.0f;
            nm[r] = 1;
            if (zin && k < jb.Nk) {
                const size_t p = (size_t)row * jb.rs + (size_t)k * jb.ks;
                nv[r] = jb.in[p * jb.SLN + z0 + lz];
                if (jb.mask) nm[r] = jb.mask[p];
            }
        }
    };
    constexpr int Q = G / 4;  // V4: slice quads per position; KC = 256 / Q positions
    const int lq = tid % Q, lkq = tid / Q;
    const bool qin = z0 + 4 * lq < jb.SLN;  // a quad lies wholly inside or outside
    float4 n4;
    uint8_t nm4;
    auto fetch4 = [&](int k0) __attribute__((always_inline)) {
        const int k = k0 + lkq;
        n4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        nm4 = 1;
        if (qin && k < jb.Nk) {
            const size_t p = (size_t)row * jb.rs + (size_t)k * jb.ks;
            n4 = *reinterpret_cast<const float4*>(jb.in + p * jb.SLN + z0 + 4 * lq);
            if (jb.mask) nm4 = jb.mask[p];
        }
    };
    // A row of Nk samples has Nk + 1 entries.  When Nk is a whole number of
    // chunks the last entry, the row total, would be a chunk of its own: it is
    // written instead with the carry of the chunk before, by the expression
    // that chunk would have evaluated (the carry plus an empty part, sample 0).
    const int kend = jb.Nk % KC == 0 ? jb.Nk : K;
    if constexpr (V4) fetch4(0);
    else fetch(0);
    for (int k0 = 0; k0 < kend; k0 += KC) {
        if constexpr (V4) {
            const float m = (float)nm4;
            float* tr = &tin[lkq][4 * lq];
            tr[0] = jb.mask ? n4.x * m : n4.x;
            tr[1] = jb.mask ? n4.y * m : n4.y;
            tr[2] = jb.mask ? n4.z * m : n4.z;
            tr[3] = jb.mask ? n4.w * m : n4.w;
            if (k0 + KC < kend) fetch4(k0 + KC);
        } else {
#pragma unroll
            for (int r = 0; r < LP; ++r) tin[r * P + lk][lz] = jb.mask ? nv[r] * (float)nm[r] : nv[r];
            if (k0 + KC < kend) fetch(k0 + KC);
        }
        __syncthreads();
        S acc;
        Lt pre[LP];
        float sv[LP];
#pragma unroll
        for (int i = 0; i < LP; ++i) {
            sv[i] = tin[part * LP + i][sz];
            pre[i] = acc.get();
            acc.add(sv[i]);
        }
        tot[part][sz] = acc.get();
        __syncthreads();
        Lt off = carry[sz];
        for (int q = 0; q < part; ++q) off = S::plus(off, tot[q][sz]);
#pragma unroll
        for (int i = 0; i < LP; ++i) tout[sz][part * LP + i] = S::entry(S::plus(off, pre[i]), sv[i]);
        __syncthreads();  // every thread has read carry before it moves on
        if (part == P - 1) {
            const Lt c = S::plus(off, acc.get());
            carry[sz] = c;
            const int zz = z0 + sz;
            if (kend < K && k0 + KC == kend && zz < SLNp) {
                const int zc = zz / jb.ZT, zt = zz - zc * jb.ZT;
                E* const dst = &base[(((size_t)zc * jb.nrows + row) * jb.ZT + zt) * K + jb.Nk];
                const E e = S::entry(S::plus(c, S::zero()), 0.0f);
                if (L.nt) store_nt(dst, e);
                else *dst = e;
            }
        }
#pragma unroll
        for (int j = 0; j < LP; ++j) {
            const int w = tid + 256 * j;
            const int z = w / KC, kk = w - z * KC;
            const int zz = z0 + z, k = k0 + kk;
            if (k < K && zz < SLNp) {
                const int zc = zz / jb.ZT, zt = zz - zc * jb.ZT;
                E* const dst = &base[(((size_t)zc * jb.nrows + row) * jb.ZT + zt) * K + k];
                if (L.nt) store_nt(dst, tout[z][kk]);
                else *dst = tout[z][kk];
            }
        }
        __syncthreads();
    }
}

// T(c) from one entry: scalar for f32/f64, (hi, lo) for the two-float table.
template <int MODE>
struct Samp;
template <>
struct Samp<SAT_F32> {
    float v;
    __device__ __forceinline__ Samp(float2 e, float t) : v(fmaf(t, e.y, e.x)) {}
    __device__ __forceinline__ float minus_prev_lane() const { return v - lane_prev(v); }
    __device__ __forceinline__ float minus(const Samp& o) const { return v - o.v; }
};
template <>
struct Samp<SAT_TWOFLOAT> {
    float h, l;
    __device__ __forceinline__ Samp(float4 e, float t) : h(e.x), l(fmaf(t, e.z, e.y)) {}
    __device__ __forceinline__ float minus_prev_lane() const {
        return (h - lane_prev(h)) + (l - lane_prev(l));
    }
    __device__ __forceinline__ float minus(const Samp& o) const { return (h - o.h) + (l - o.l); }
};
template <>
struct Samp<SAT_F64> {
    double v;
    __device__ __forceinline__ Samp(double2 e, float t) : v(fma((double)t, e.y, e.x)) {}
    __device__ __forceinline__ float minus_prev_lane() const { return (float)(v - lane_prev(v)); }
    __device__ __forceinline__ float minus(const Samp& o) const { return (float)(v - o.v); }
};

// ------------------------------------------------------------------ projector
// Branchless DD projector.  One wave = V views x one slice chunk x one group of
// 63 detectors; lane i owns detector boundary ray b = 63*group + i, so every
// boundary is intersected with each slab once and its table sample T_b is
// handed to the next lane through DPP: detector u = b - 1 accumulates
// (T_b - T_{b-1}) / (r_b - r_{b-1}), the reference's
// (T(realR) - T(realL)) / intersectLength (MultiSlicesProjGPU.cu:160-171,
// 183-194), since neighbouring detectors share the boundary ray bit for bit.
// The V views of a wave are consecutive views of the subset: at each slab they
// read overlapping stretches of the same table rows, which then come from L1.
//
// Slab passes: a launch covers slabs [ii0, ii1) only, so every block resident on
// an XCD reads the same band of table rows and the band stays in that XCD's L2
// for all views; pass sums accumulate in `out` (first pass writes, last pass
// applies the epilogue), in a fixed order.
struct ProjDev {
    GeomDev g;
    Subset sub;
    int nzc, nug, nvg, blk0, ii0, ii1, first, last, epi;
    int zt, vb;  // k_proj_rows: the tables' slice chunk and the views of a block
    const float2* cossin;
    const float* bx;
    const float* by;
    const float* xds;
    const float* yds;
    const void* satX;
    const void* satY;
    float* out;
    const float* meas;
    const float* rowSum;
};
struct ProjLaunch {
    int njobs, nwork;
    ProjDev j[kBatch];
};

// Waves per SIMD the projector and backprojector are compiled for (the
// compiler's register budget).  Measured on MI355X at 512x512x64, 2304 views
// (two-float tables): the projector keeps a view's ZT table loads in flight
// together in the registers of 5 waves (29.9 ms; 6 waves serialise them,
// 39 ms); the backprojector's gathers contend in the texture address unit,
// where 8 waves with few loads in flight each beat 7 waves with many (34.1 vs
// 38.9 ms).
#ifndef DDF_PROJ_WAVES
#define DDF_PROJ_WAVES 5
#endif
#ifndef DDF_BACK_WAVES
#define DDF_BACK_WAVES 8
#endif


// Writes one detector's NQ sums (slices z0..) of subset view j: plain (the
// operator; later slab passes add to the earlier ones) or, on the last pass,
// the OS-SART ratio (b - p) / rowSum (OSSART_AAPM.m:81-88).
template <int NQ>
__device__ __forceinline__ void proj_store_view(const ProjDev& J, const GeomDev& g,
                                                const Subset& sub, int j, int u, int z0,
                                                const float (&acc)[NQ], float cval) {
    const bool ratio = J.epi == PROJ_RATIO && J.last;
    float* __restrict__ out = J.out;
    const int vg = sub.first + j * sub.stride;
    const size_t o0 = ((size_t)j * g.DNU + u) * g.SLN + z0;
    const float w = ratio ? J.rowSum[(size_t)j * g.DNU + u] : 0.0f;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int z = z0 + q;
        if (z < g.SLN) {
            float p = acc[q] * cval;
            if (!J.first) p = p + out[o0 + q];
            if (!ratio) {
                out[o0 + q] = p;
            } else {
                const float b = J.meas[((size_t)vg * g.DNU + u) * g.SLN + z];
                out[o0 + q] = (w != 0.0f) ? (b - p) / w : p;
            }
        }
    }
}

// The ZT sums of the V views of view group jg.
template <int ZT, int V>
__device__ __forceinline__ void proj_store(const ProjDev& J, const GeomDev& g, const Subset& sub,
                                           int jg, int u, int z0, const float (&acc)[V][ZT],
                                           const float (&cval)[V]) {
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int j = jg * V + v;
        if (j >= sub.count) break;
        proj_store_view<ZT>(J, g, sub, j, u, z0, acc[v], cval[v]);
    }
}

// Continuous factor of detector u in a view with angle (c, s): dx / the
// |normalize(det - src)| component along the slab normal
// (MultiSlicesProjGPU.cu:139,158,181).
__device__ __forceinline__ float proj_cval(const GeomDev& g, float c, float s, bool xslab,
                                           float xdu, float ydu) {
    const float srcx = g.x0 * c - g.y0 * s, srcy = g.x0 * s + g.y0 * c;
    const float ddx = (xdu * c - ydu * s) - srcx;
    const float ddy = (xdu * s + ydu * c) - srcy;
    const float nrm = sqrtf(fmaf(ddx, ddx, ddy * ddy));
    return g.dx * nrm / fabsf(xslab ? ddx : ddy);
}

template <int MODE, int ZT, int V>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(DDF_PROJ_WAVES, DDF_PROJ_WAVES)))
void k_proj(const ProjLaunch L) {
    using E = typename Sat<MODE>::E;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= L.nwork) return;  // whole wave
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (lin >= L.j[t].blk0) ji = t;
    const ProjDev& J = L.j[ji];
    const GeomDev g = J.g;
    const Subset sub = J.sub;
    const int local = lin - J.blk0;
    const int ug = local % J.nug;
    const int rest = local / J.nug;
    const int jg = rest % J.nvg;
    const int zc = rest / J.nvg;
    // chunk zc of both tables: satX rows are y (prefix along x), satY rows are x
    const E* baseX = reinterpret_cast<const E*>(J.satX) + (size_t)zc * g.YN * (size_t)ZT * (g.XN + 1);
    const E* baseY = reinterpret_cast<const E*>(J.satY) + (size_t)zc * g.XN * (size_t)ZT * (g.YN + 1);

    const int lane = threadIdx.x;
    const int bnd = min(ug * 63 + lane, g.DNU);
    const int u = ug * 63 + lane - 1;
    const bool act = lane >= 1 && u < g.DNU;
    const int uu = min(max(u, 0), g.DNU - 1);
    const int z0 = zc * ZT;
    const float bxb = J.bx[bnd], byb = J.by[bnd], xdu = J.xds[uu], ydu = J.yds[uu];

    // Per view: the boundary ray of this lane and the continuous factor; the
    // slab orientation of view v is bit v of xs (a wave-uniform choice, so the
    // per-view table selection below stays scalar).
    float pa[V], pf[V], pbb[V], cval[V];
    unsigned xs = 0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int j = min(jg * V + v, sub.count - 1);
        const int vg = sub.first + j * sub.stride;
        const float2 cs = J.cossin[(size_t)vg * g.SLN + min(z0, g.SLN - 1)];
        const ProjBoundary pb = proj_boundary(g.x0, g.y0, cs.x, cs.y, bxb, byb);
        cval[v] = proj_cval(g, cs.x, cs.y, pb.xslab, xdu, ydu);
        pa[v] = pb.a;
        pf[v] = pb.f;
        pbb[v] = pb.b;
        if (proj_xslab(cs.x, cs.y)) xs |= 1u << v;
    }

    float acc[V][ZT];
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int q = 0; q < ZT; ++q) acc[v][q] = 0.0f;
    int iiEnd = 0;
    bool sameEnd = true;  // every view of the wave has the same slab count
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int n = ((xs >> v) & 1u) ? g.XN : g.YN;
        sameEnd = sameEnd && n == (((xs & 1u) ? g.XN : g.YN));
        iiEnd = max(iiEnd, min(J.ii1, n));
    }
    // Few loads per slab (ZT x V <= 4) leave a wave waiting for each slab's
    // gathers: there the entries of slab ii + 1 are loaded before the sums of
    // slab ii (the same arithmetic in the same order, so the same results).
    if constexpr (ZT * V <= 4) {
        if (sameEnd && J.ii0 < iiEnd) {
            E e[V][ZT];
            float tw[V], il[V];
            auto fetch = [&](int ii, E (&ee)[V][ZT], float (&t)[V], float (&inv)[V])
                __attribute__((always_inline)) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const bool xv = (xs >> v) & 1u;
                    const int Nk = xv ? g.YN : g.XN;
                    const float ocS = xv ? g.ocxP : g.ocyP, ocK = xv ? g.ocyP : g.ocxP;
                    const int K = Nk + 1;
                    const float r = slab_r(slab_obj(ii, ocS, g.dx), pa[v], pf[v], pbb[v]);
                    const Fetch f = fetch_pos<E>(slab_coord(r, g.invdx, ocK), Nk);
                    inv[v] = __builtin_amdgcn_rcpf(r - lane_prev(r));
                    t[v] = f.t;
                    const E* row = (xv ? baseY : baseX) + (size_t)ii * ZT * K;
#pragma unroll
                    for (int q = 0; q < ZT; ++q) ee[v][q] = entry_at(row + (size_t)q * K, f.off);
                }
            };
            fetch(J.ii0, e, tw, il);
            for (int ii = J.ii0; ii < iiEnd; ++ii) {
                E en[V][ZT];
                float tn[V], iln[V];
                fetch(min(ii + 1, iiEnd - 1), en, tn, iln);  // the last slab loads itself again
#pragma unroll
                for (int v = 0; v < V; ++v)
#pragma unroll
                    for (int q = 0; q < ZT; ++q) {
                        const Samp<MODE> T(e[v][q], tw[v]);
                        acc[v][q] = fmaf(T.minus_prev_lane(), il[v], acc[v][q]);
                    }
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    tw[v] = tn[v];
                    il[v] = iln[v];
#pragma unroll
                    for (int q = 0; q < ZT; ++q) e[v][q] = en[v][q];
                }
            }
            if (!act) return;
            proj_store<ZT, V>(J, g, sub, jg, u, z0, acc, cval);
            return;
        }
    }
    for (int ii = J.ii0; ii < iiEnd; ++ii) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const bool xv = (xs >> v) & 1u;
            if (ii < (xv ? g.XN : g.YN)) {  // uniform across the wave
                const int Nk = xv ? g.YN : g.XN;
                const float ocS = xv ? g.ocxP : g.ocyP, ocK = xv ? g.ocyP : g.ocxP;
                const int K = Nk + 1;
                const float r = slab_r(slab_obj(ii, ocS, g.dx), pa[v], pf[v], pbb[v]);
                const Fetch f = fetch_pos<E>(slab_coord(r, g.invdx, ocK), Nk);
                const float invlen = __builtin_amdgcn_rcpf(r - lane_prev(r));
                E e[ZT];
                const E* row = (xv ? baseY : baseX) + (size_t)ii * ZT * K;
#pragma unroll
                for (int q = 0; q < ZT; ++q) e[q] = entry_at(row + (size_t)q * K, f.off);
#pragma unroll
                for (int q = 0; q < ZT; ++q) {
                    const Samp<MODE> T(e[q], f.t);
                    acc[v][q] = fmaf(T.minus_prev_lane(), invlen, acc[v][q]);
                }
            }
        }
    }
    if (!act) return;
    proj_store<ZT, V>(J, g, sub, jg, u, z0, acc, cval);
}

// ------------------------------------------------------------------ backprojector
// Voxel-driven DD backprojector: a 256-thread block is a 16x16 voxel tile of
// one slice chunk; each lane loops over the subset's views, computes the
// voxel footprint on the detector once for the chunk and accumulates
// (T(idxR) - T(idxL)) * cosVal for its ZT slices.  Masked voxels skip the
// work and write 0 (or leave the iterate unchanged in the OS-SART epilogues).
struct BackDev {
    GeomDev g;
    Subset sub;
    int nzc, ntx, nty, blk0, epi;
    int tabViews;  // views per slice chunk of the table P (>= sub.count)
    int accum;     // 1: the sums continue the values in vol (a later view pass)
    const float2* src;  // source positions [view][slice]
    const uint8_t* mask;
    const void* P;
    float* vol;
    const float* colSum;
    float* xprev;
    const float* momTab;
    const int* counter;
};
struct BackLaunch {
    int njobs, nwork;
    BackDev j[kBatch];
};

// Writes one voxel column's ZT sums: plain (the operator) or the OS-SART
// update x += sum / colSum with the FISTA extrapolation on a subset's last
// launch (OSSART_AAPM.m:98-115).
__device__ __forceinline__ void back_store(const BackDev& J, const GeomDev& g, int x, int y, int z0,
                                           const float* acc, int ZT) {
    const size_t pix = (size_t)y * g.XN + x;
    const size_t o0 = pix * g.SLN + z0;
    float* vol = J.vol;
    if (J.epi == BACK_PLAIN) {
        for (int q = 0; q < ZT; ++q)
            if (z0 + q < g.SLN) vol[o0 + q] = acc[q];
        return;
    }
    const float w = J.colSum[pix];
    const bool fista = J.epi == BACK_UPDATE_FISTA;
    const float mom = fista ? J.momTab[*J.counter] : 0.0f;
    for (int q = 0; q < ZT; ++q) {
        if (z0 + q < g.SLN) {
            const float upd = (w != 0.0f) ? acc[q] / w : acc[q];
            const float xn = vol[o0 + q] + upd;
            if (!fista) {
                vol[o0 + q] = xn;
            } else {
                const float xo = xn + mom * (xn - J.xprev[o0 + q]);
                vol[o0 + q] = xo;
                J.xprev[o0 + q] = xo;
            }
        }
    }
}

template <int MODE, int ZT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DDF_BACK_WAVES)))
void k_back(const BackLaunch L) {
    using E = typename Sat<MODE>::E;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= L.nwork) return;
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (lin >= L.j[t].blk0) ji = t;
    const BackDev& J = L.j[ji];
    const GeomDev& g = J.g;
    const int local = lin - J.blk0;
    const int tile = local % (J.ntx * J.nty);
    const int zc = local / (J.ntx * J.nty);
    const int x = (tile % J.ntx) * 16 + (threadIdx.x & 15);
    const int y = (tile / J.ntx) * 16 + (threadIdx.x >> 4);
    const bool inb = x < g.XN && y < g.YN;
    const int m = inb ? J.mask[(size_t)y * g.XN + x] : 0;
    const int z0 = zc * ZT;
    const int zg = min(z0, g.SLN - 1);
    const int DNU = g.DNU;
    const int K = DNU + 1;
    const size_t viewStride = (size_t)ZT * K;
    const E* Pc = reinterpret_cast<const E*>(J.P) + (size_t)zc * J.tabViews * viewStride;

    float acc[ZT];
#pragma unroll
    for (int q = 0; q < ZT; ++q)
        acc[q] = (J.accum && inb && z0 + q < g.SLN) ? J.vol[((size_t)y * g.XN + x) * g.SLN + z0 + q]
                                                    : 0.0f;
    if (m != 0) {
        const float vx = ((float)x - g.ocxB) * g.dx;
        const float vy = ((float)y - g.ocyB) * g.dx;
        for (int j = 0; j < J.sub.count; ++j) {
            const int vg = J.sub.first + j * J.sub.stride;
            const float2 sp = J.src[(size_t)vg * g.SLN + zg];
            const BackSample bs =
                back_sample(vx, vy, sp.x, sp.y, g.inv_sid, g.invdb, g.detCntIdx, g.dx, g.backMid);
            const Fetch fL = fetch_pos<E>(bs.idxL, DNU), fR = fetch_pos<E>(bs.idxR, DNU);
            // dx / ddv * sqrt(l2) with sqrt(l2) = 1 / l  (MultiSlicesBackGPU.cu:173)
            const float cosVal = g.dx * __builtin_amdgcn_rcpf(bs.ddv * bs.l);
            const E* row = Pc + (size_t)j * viewStride;
#pragma unroll
            for (int q = 0; q < ZT; ++q) {
                const Samp<MODE> TL(entry_at(row + (size_t)q * K, fL.off), fL.t);
                const Samp<MODE> TR(entry_at(row + (size_t)q * K, fR.off), fR.t);
                acc[q] = fmaf(TR.minus(TL), cosVal, acc[q]);
            }
        }
    }
    if (!inb) return;
    back_store(J, g, x, y, z0, acc, ZT);
}

// Staged backprojector: the sums of k_back, with the table reads served from
// LDS.  For each group of NV consecutive views the block copies, per (view,
// slice), the run of detector entries the tile's footprint covers (at most
// 64 entries) into a 64-entry LDS row with one 64-lane global-to-LDS load,
// and the lanes gather from LDS instead of through the texture-address path
// (one address per lane per slice, the bound of k_back).  Two buffers
// alternate: the copy of the next group is in flight while the current group
// is summed.  A view's run spans the footprints of the tile's four corner
// voxels, one entry wider on each side (a wider footprint gets the 64 entries
// centred on it), which every wave computes the same way (no exchange
// needed); a lane whose entries fall outside the run reads the table in
// global memory, so the run decides speed only.  Arithmetic and summation
// order are those of k_back, so the results are bit-identical.
// Staging: one 64-lane global-to-LDS copy of 16 bytes per lane (one two-float
// or f64 entry, two f32 entries) from a table row into LDS.  LDS operands are
// typed as LDS pointers throughout, so the compiler tracks each staging buffer
// as its own object and leaves no generic-to-LDS conversion to check.
#define DDF_LDS __attribute__((address_space(3)))
template <class E>
__device__ __forceinline__ void stage_row(const E* src, DDF_LDS void* dst, int nEntries, int lane) {
    constexpr int PER = 16 / (int)sizeof(E);  // entries one lane copies
    if (lane * PER < nEntries)
        __builtin_amdgcn_global_load_lds(static_cast<const void*>(src + lane * PER), dst, 16, 0, 0);
}
// LDS entry reads as native vectors.  A two-float entry is read as one 16-byte
// vector: the compiler keeps that a single ds_read_b128, where the float4 of
// the global path narrows to a 12-byte ds_read_b96.  Measured on MI355X at
// 512x512x64, 2304 views: backprojector 20.3 ms with 16-byte reads, 25.4 ms
// with 12-byte ones; reading the 12 bytes used as ds_read_b64 + ds_read_b32:
// projector 15.6 -> 18.2 ms, backprojector 19.2 -> 21.2 ms.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 lds_entry(const DDF_LDS float2* p) {
    const f2v v = *(const DDF_LDS f2v*)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 lds_entry(const DDF_LDS float4* p) {
    const f4v v = *(const DDF_LDS f4v*)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ double2 lds_entry(const DDF_LDS double2* p) {
    const d2v v = *(const DDF_LDS d2v*)p;
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
// Wave-uniform element of a table the kernel only reads, through the scalar
// cache (s_load, counted in lgkmcnt): waiting for it does not wait for the
// global-to-LDS copies still in flight, which a vector load's vmcnt would.
__device__ __forceinline__ float2 uniform_ld(const float2* p, size_t i) {
    const f2v v = ((const __attribute__((address_space(4))) f2v*)p)[i];
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ int quad_min(int v) {
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));  // quad_perm [1,0,3,2]
    return min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));  // quad_perm [2,3,0,1]
}
__device__ __forceinline__ int quad_max(int v) {
    v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));
    return max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));
}

// Staged rows hold 64 entries: 1 KiB of two-float or f64 entries, 512 bytes of
// f32 ones.  Blocks per CU the two staging buffers (2 * NV * ZT rows) leave
// room for in 160 KiB of LDS, each block putting one wave on each SIMD, at
// most 8 (6 for f32 rows, whose small buffers would otherwise leave the
// compiler a 64-register budget).  Measured on MI355X at 512x512x64, 2304
// views, 16-slice chunks, one view per group: f32 15.7 ms with 128-entry rows
// (5 waves), 14.9 ms with 64-entry rows (6 waves); two-float 16.9 ms with
// 64-entry rows, 18.6 ms with 48 and 28.6 ms with 32 (tiles whose detector
// footprint is wider than the row read their entries from global memory).
constexpr int kBackWent = 64;
constexpr int back_lds_waves(int MODE, int NV, int ZT) {
    const int rowBytes = kBackWent * (MODE == SAT_F32 ? 8 : 16);
    const int w = 160 * 1024 / (2 * NV * ZT * rowBytes);
    const int cap = MODE == SAT_F32 ? 6 : 8;
    return w < cap ? w : cap;
}

template <int MODE, int ZT, int NV>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(back_lds_waves(MODE, NV, ZT))))
void k_back_lds(const BackLaunch L) {
    using E = typename Sat<MODE>::E;
    constexpr int WENT = kBackWent;  // entries per staged row
    constexpr int ROWS = NV * ZT;
    static_assert(NV >= 1 && NV <= 16, "four corner lanes per view");
    __shared__ __attribute__((aligned(16))) E bufA[ROWS * WENT];
    __shared__ __attribute__((aligned(16))) E bufB[ROWS * WENT];
    using LE = DDF_LDS E;
    LE* const ldsA = (LE*)bufA;
    LE* const ldsB = (LE*)bufB;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= L.nwork) return;
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (lin >= L.j[t].blk0) ji = t;
    const BackDev& J = L.j[ji];
    const GeomDev& g = J.g;
    const int local = lin - J.blk0;
    const int tile = local % (J.ntx * J.nty);
    const int zc = local / (J.ntx * J.nty);
    const int tx0 = (tile % J.ntx) * 16, ty0 = (tile / J.ntx) * 16;
    // Lane -> voxel: a wave is an 8x8 voxel patch and each 16 lanes a 4x4 one,
    // so neighbouring lanes read nearby detector entries in every view
    // direction (a row of 16 voxels spans ~22 entries when rays run along y).
    // Measured on MI355X at 512x512x64, 2304 views, against 16x4 row-major
    // waves: two-float 16.9 -> 15.2 ms, f32 14.7 -> 11.8, f64 19.9 -> 18.9
    // (fewer LDS bank conflicts; 4x4 patches in 16x4 waves do the same).
    const int x = tx0 + 8 * ((threadIdx.x >> 6) & 1) + 4 * ((threadIdx.x >> 4) & 1) + (threadIdx.x & 3);
    const int y = ty0 + 8 * (threadIdx.x >> 7) + 4 * ((threadIdx.x >> 5) & 1) + ((threadIdx.x >> 2) & 3);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const bool inb = x < g.XN && y < g.YN;
    const int m = inb ? J.mask[(size_t)y * g.XN + x] : 0;
    const int z0 = zc * ZT;
    const int zg = min(z0, g.SLN - 1);
    const int DNU = g.DNU;
    const int K = DNU + 1;
    const int count = J.sub.count;
    const size_t viewStride = (size_t)ZT * K;
    const E* Pc = reinterpret_cast<const E*>(J.P) + (size_t)zc * J.tabViews * viewStride;

    float acc[ZT];
#pragma unroll
    for (int q = 0; q < ZT; ++q)
        acc[q] = (J.accum && inb && z0 + q < g.SLN) ? J.vol[((size_t)y * g.XN + x) * g.SLN + z0 + q]
                                                    : 0.0f;
    // Any unmasked voxel in the tile?  (Reduced through the first staging
    // buffer, so the block's LDS is exactly the two buffers.)
    {
        const unsigned long long b = __ballot(m != 0);
        if (lane == 0) reinterpret_cast<int*>(bufA)[wave] = b != 0ull;
    }
    __syncthreads();
    const int* fl = reinterpret_cast<const int*>(bufA);
    const bool any = (fl[0] | fl[1] | fl[2] | fl[3]) != 0;
    __syncthreads();
    if (any && count > 0) {
        // Window starts of 16 consecutive views v0.., one per quad of lanes
        // (lane = 4*(view - v0) + corner of the tile).
        const int cx = (lane & 1) ? min(tx0 + 15, g.XN - 1) : tx0;
        const int cy = (lane & 2) ? min(ty0 + 15, g.YN - 1) : ty0;
        const float cvx = ((float)cx - g.ocxB) * g.dx, cvy = ((float)cy - g.ocyB) * g.dx;
        int wl = 0, wn = 0;
        auto windows16 = [&](int v0) __attribute__((always_inline)) {
            const int j = min(v0 + (lane >> 2), count - 1);
            const int vg = J.sub.first + j * J.sub.stride;
            const float2 sp = J.src[(size_t)vg * g.SLN + zg];
            const BackSample bs = back_sample(cvx, cvy, sp.x, sp.y, g.inv_sid, g.invdb,
                                              g.detCntIdx, g.dx, g.backMid);
            const int lo = quad_min(texel(fminf(bs.idxL, bs.idxR), DNU));
            const int hi = quad_max(texel(fmaxf(bs.idxL, bs.idxR), DNU));
            const int s = max(lo - 1, 0), e = min(hi + 1, DNU);
            if (e - s < WENT) {
                wl = s;
                wn = e - s + 1;
            } else {
                wl = max(0, min(((lo + hi) >> 1) - WENT / 2, K - WENT));
                wn = min(K - wl, WENT);
            }
        };
        auto window = [&](int grp, int* w0, int* n0) __attribute__((always_inline)) {
            const int j0 = grp * NV;  // NV divides 16: a group never straddles two sets
            if ((j0 & 15) == 0) windows16(j0);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                w0[v] = __builtin_amdgcn_readlane(wl, 4 * ((j0 + v) & 15));
                n0[v] = __builtin_amdgcn_readlane(wn, 4 * ((j0 + v) & 15));
            }
        };
        // Row (view v, slice q) is copied by wave (v*ZT + q) mod 4: its WENT
        // entries, or the rest of the table row if shorter.
        auto stage = [&](int grp, LE* buf, const int* w0, const int* n0) __attribute__((always_inline)) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int j = grp * NV + v;
                if (j >= count) break;
                const E* src = Pc + (size_t)j * viewStride + w0[v];
                const int nrow = n0[v];
                if constexpr (ZT % 4 == 0) {
#pragma unroll
                    for (int i = 0; i < ZT / 4; ++i) {
                        const int q = 4 * i + wave;
                        stage_row(src + (size_t)q * K, (DDF_LDS void*)(buf + (v * ZT + q) * WENT),
                                      nrow, lane);
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < ZT; ++q)
                        if (((v * ZT + q) & 3) == wave)
                            stage_row(src + (size_t)q * K,
                                          (DDF_LDS void*)(buf + (v * ZT + q) * WENT), nrow,
                                          lane);
                }
            }
        };
        const float vx = ((float)x - g.ocxB) * g.dx;
        const float vy = ((float)y - g.ocyB) * g.dx;
        auto sum = [&](int grp, const LE* buf, const int* w0, const int* n0) __attribute__((always_inline)) {
            if (m == 0) return;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int j = grp * NV + v;
                if (j >= count) break;
                const int vg = J.sub.first + j * J.sub.stride;
                const float2 sp = uniform_ld(J.src, (size_t)vg * g.SLN + zg);
                const BackSample bs = back_sample(vx, vy, sp.x, sp.y, g.inv_sid, g.invdb,
                                                  g.detCntIdx, g.dx, g.backMid);
                const Fetch fL = fetch_pos<E>(bs.idxL, DNU), fR = fetch_pos<E>(bs.idxR, DNU);
                const float cosVal = g.dx * __builtin_amdgcn_rcpf(bs.ddv * bs.l);
                const int iL = fL.i0 - w0[v];
                const int iR = fR.i0 - w0[v];
                if ((unsigned)iL < (unsigned)n0[v] && (unsigned)iR < (unsigned)n0[v]) {
                    const LE* lrow = buf + v * ZT * WENT;
#pragma unroll
                    for (int q = 0; q < ZT; ++q) {
                        const Samp<MODE> TL(lds_entry(lrow + q * WENT + iL), fL.t);
                        const Samp<MODE> TR(lds_entry(lrow + q * WENT + iR), fR.t);
                        acc[q] = fmaf(TR.minus(TL), cosVal, acc[q]);
                    }
                } else {
                    const E* row = Pc + (size_t)j * viewStride;
#pragma unroll
                    for (int q = 0; q < ZT; ++q) {
                        const Samp<MODE> TL(entry_at(row + (size_t)q * K, fL.off), fL.t);
                        const Samp<MODE> TR(entry_at(row + (size_t)q * K, fR.off), fR.t);
                        acc[q] = fmaf(TR.minus(TL), cosVal, acc[q]);
                    }
                }
            }
        };
        int wA[NV], wB[NV], nA[NV], nB[NV];
        const int ng = (count + NV - 1) / NV;
        window(0, wA, nA);
        stage(0, ldsA, wA, nA);
        wait_vmem();
        __syncthreads();
        for (int grp = 0; grp < ng; grp += 2) {
            if (grp + 1 < ng) {
                window(grp + 1, wB, nB);
                stage(grp + 1, ldsB, wB, nB);
            }
            sum(grp, ldsA, wA, nA);
            wait_vmem();
            __syncthreads();
            if (grp + 1 >= ng) break;
            if (grp + 2 < ng) {
                window(grp + 2, wA, nA);
                stage(grp + 2, ldsA, wA, nA);
            }
            sum(grp + 1, ldsB, wB, nB);
            wait_vmem();
            __syncthreads();
        }
    }
    if (!inb) return;
    back_store(J, g, x, y, z0, acc, ZT);
}

// Staged projector: the sums of k_proj, with the table reads served from LDS.
// A 256-thread block is four waves = four consecutive view groups (4V views) of
// one detector group and slice chunk; at each slab their boundary rays cross
// the same stretch of one table row.  For groups of S slabs the block copies,
// per (slab, slice), the WP-entry run of the row centred on the crossing of
// one ray (the middle boundary of the block's middle view, which every wave
// evaluates the same way) into LDS with 64-lane global-to-LDS loads, two
// buffers alternating, and the lanes gather from LDS.  Views whose slab
// orientation differs from the middle view's, and lanes whose entry falls
// outside the run, read the table in global memory.  Arithmetic and summation
// order are those of k_proj, so the results are bit-identical.
// Register budget: 5 waves per SIMD (the projector's), fewer where the ZT x V
// accumulators and loads would spill.  Measured at 512x512x64, 2304 views,
// chunk 8 x 4 views per wave: two-float tables 16.7 ms at 4 waves vs 22.6 ms at
// 5 (spilling); f32 tables 10.5 ms at 5 waves (a 16-byte spill) vs 11.5 at 4.
constexpr int proj_lds_waves(int MODE, int ZT, int V) {
    return ZT * V >= 64 ? 2 : (ZT * V >= 32 && MODE != SAT_F32) ? 4 : 5;
}
template <int MODE, int ZT, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(proj_lds_waves(MODE, ZT, V))))
void k_proj_lds(const ProjLaunch L) {
    using E = typename Sat<MODE>::E;
    constexpr int WP = 128;                          // entries per staged run
    constexpr int PER = 16 / (int)sizeof(E);         // entries per lane per copy
    constexpr int CP = WP / (64 * PER);              // 64-lane copies per run
    constexpr int S = ZT >= 8 ? 1 : 8 / ZT;          // slabs per group
    constexpr int ROWS = S * ZT;
    // Two buffers as separate LDS objects: the compiler then knows that reads of
    // one do not alias the global-to-LDS copies in flight into the other and
    // puts no wait for those copies in front of them.
    __shared__ __attribute__((aligned(16))) E bufA[ROWS * WP];
    __shared__ __attribute__((aligned(16))) E bufB[ROWS * WP];
    using LE = DDF_LDS E;
    LE* const ldsA = (LE*)bufA;
    LE* const ldsB = (LE*)bufB;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= L.nwork) return;  // whole block
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (lin >= L.j[t].blk0) ji = t;
    const ProjDev& J = L.j[ji];
    const GeomDev g = J.g;
    const Subset sub = J.sub;
    const int local = lin - J.blk0;
    const int ug = local % J.nug;
    const int rest = local / J.nug;
    const int jb = rest % J.nvg;  // block of four view groups
    const int zc = rest / J.nvg;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int jg = jb * 4 + wave;
    const E* baseX = reinterpret_cast<const E*>(J.satX) + (size_t)zc * g.YN * (size_t)ZT * (g.XN + 1);
    const E* baseY = reinterpret_cast<const E*>(J.satY) + (size_t)zc * g.XN * (size_t)ZT * (g.YN + 1);

    const int bnd = min(ug * 63 + lane, g.DNU);
    const int u = ug * 63 + lane - 1;
    const bool act = lane >= 1 && u < g.DNU;
    const int uu = min(max(u, 0), g.DNU - 1);
    const int z0 = zc * ZT;
    const int zs = min(z0, g.SLN - 1);
    const float bxb = J.bx[bnd], byb = J.by[bnd], xdu = J.xds[uu], ydu = J.yds[uu];

    float pa[V], pf[V], pbb[V], cval[V];
    unsigned xs = 0;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        const int j = min(jg * V + v, sub.count - 1);
        const int vg = sub.first + j * sub.stride;
        const float2 cs = J.cossin[(size_t)vg * g.SLN + zs];
        const ProjBoundary pb = proj_boundary(g.x0, g.y0, cs.x, cs.y, bxb, byb);
        cval[v] = proj_cval(g, cs.x, cs.y, pb.xslab, xdu, ydu);
        pa[v] = pb.a;
        pf[v] = pb.f;
        pbb[v] = pb.b;
        if (proj_xslab(cs.x, cs.y)) xs |= 1u << v;
    }

    // The block's staged table and its window ray.
    const int jm = min(jb * 4 * V + 2 * V, sub.count - 1);
    const float2 csm = J.cossin[(size_t)(sub.first + jm * sub.stride) * g.SLN + zs];
    const bool bxs = proj_xslab(csm.x, csm.y);
    const int bm = min(ug * 63 + 32, g.DNU);
    const ProjBoundary pm = proj_boundary(g.x0, g.y0, csm.x, csm.y, J.bx[bm], J.by[bm]);
    const int NkB = bxs ? g.YN : g.XN, KB = NkB + 1, NsB = bxs ? g.XN : g.YN;
    const float ocSB = bxs ? g.ocxP : g.ocyP, ocKB = bxs ? g.ocyP : g.ocxP;
    const E* tabB = bxs ? baseY : baseX;
    const unsigned xsB = bxs ? 0xFu : 0u;

    float acc[V][ZT];
#pragma unroll
    for (int v = 0; v < V; ++v)
#pragma unroll
        for (int q = 0; q < ZT; ++q) acc[v][q] = 0.0f;

    const int ii0 = J.ii0;
    const int iiE = min(J.ii1, max(g.XN, g.YN));  // block-uniform
    const int ng = (iiE - ii0 + S - 1) / S;
    auto stage = [&](int grp, LE* buf, int* w0) __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ii = ii0 + grp * S + s;
            const float r = slab_r(slab_obj(ii, ocSB, g.dx), pm.a, pm.f, pm.b);
            const int c = __builtin_amdgcn_readfirstlane(texel(slab_coord(r, g.invdx, ocKB), NkB));
            w0[s] = max(0, min(c - WP / 2, KB - WP));
            if (ii >= iiE || ii >= NsB) continue;
#pragma unroll
            for (int q = 0; q < ZT; ++q)
#pragma unroll
                for (int cp = 0; cp < CP; ++cp) {
                    if ((((s * ZT + q) * CP + cp) & 3) != wave) continue;
                    const int e0 = w0[s] + cp * 64 * PER;
                    stage_row(tabB + ((size_t)ii * ZT + q) * KB + e0,
                                  (DDF_LDS void*)&buf[(s * ZT + q) * WP + cp * 64 * PER], KB - e0,
                                  lane);
                }
        }
    };
    auto sum = [&](int grp, const LE* buf, const int* w0) __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int ii = ii0 + grp * S + s;
            if (ii >= iiE) break;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const bool xv = (xs >> v) & 1u;
                if (ii < (xv ? g.XN : g.YN)) {  // uniform across the wave
                    const int Nk = xv ? g.YN : g.XN;
                    const float ocS = xv ? g.ocxP : g.ocyP, ocK = xv ? g.ocyP : g.ocxP;
                    const int K = Nk + 1;
                    const float r = slab_r(slab_obj(ii, ocS, g.dx), pa[v], pf[v], pbb[v]);
                    const Fetch f = fetch_pos<E>(slab_coord(r, g.invdx, ocK), Nk);
                    // The neighbour difference is pinned ahead of the fallback
                    // branch, where the DPP read fuses into the subtraction.
                    float dr = r - lane_prev(r);
                    asm volatile("" : "+v"(dr));
                    const float invlen = __builtin_amdgcn_rcpf(dr);
                    const int li = f.i0 - w0[s];
                    // Every lane reads the run (index clamped into it); lanes outside
                    // the run, and views of the other slab orientation, then replace
                    // their entries with global reads in a branch (skipped when no
                    // lane needs it) that waits for its own loads, so the common path
                    // waits for no vector memory (copies of the next group stay in
                    // flight).  The sums run after the branch with every lane active,
                    // as the DPP neighbour read needs.
                    const bool inw = ((xs ^ xsB) >> v & 1u) == 0u && (unsigned)li < (unsigned)WP;
                    const int lc = min(max(li, 0), WP - 1);
                    E e[ZT];
#pragma unroll
                    for (int q = 0; q < ZT; ++q) e[q] = lds_entry(&buf[(s * ZT + q) * WP + lc]);
                    if (!inw) {
                        const E* row = (xv ? baseY : baseX) + (size_t)ii * ZT * K;
#pragma unroll
                        for (int q = 0; q < ZT; ++q) e[q] = entry_at(row + (size_t)q * K, f.off);
                        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), inside the branch
                    }
#pragma unroll
                    for (int q = 0; q < ZT; ++q) {
                        const Samp<MODE> T(e[q], f.t);
                        acc[v][q] = fmaf(T.minus_prev_lane(), invlen, acc[v][q]);
                    }
                }
            }
        }
    };
    int wA[S], wB[S];
    if (ng > 0) {
        stage(0, ldsA, wA);
        wait_vmem();
        __syncthreads();
    }
    for (int grp = 0; grp < ng; grp += 2) {
        if (grp + 1 < ng) stage(grp + 1, ldsB, wB);
        sum(grp, ldsA, wA);
        wait_vmem();
        __syncthreads();
        if (grp + 1 >= ng) break;
        if (grp + 2 < ng) stage(grp + 2, ldsA, wA);
        sum(grp + 1, ldsB, wB);
        wait_vmem();
        __syncthreads();
    }
    if (!act) return;
    proj_store<ZT, V>(J, g, sub, jg, u, z0, acc, cval);
}

// Whole-row projector for view subsets whose neighbouring views are far apart
// (OS-SART subsets k, k + numOS, ...): no two views of a block cross the same
// short stretch of a row, so k_proj_lds's staged runs do not serve them, and
// the per-lane gathers of k_proj keep the texture address unit saturated.
// Here a block of NW waves owns up to VB views of the subset, all their
// detector boundaries and ZS slices of a table chunk, and copies, slab after
// slab, the ZS whole table rows of that slab into LDS (global-to-LDS copies,
// two buffers alternating: the next slab's rows are in flight while the
// current one is summed); every boundary then reads its entries from LDS,
// where a slab's rows serve all the block's views.  A block's views share one
// slab orientation (one table): the subset is split into runs of consecutive
// views of one orientation and each run into blocks of VB views, so no block
// sweeps the slabs twice.  Tasks are (view, group of 63 detectors): slot s of
// wave w is task NW*s + w, T slots per wave.  Each task's terms and their
// order are those of k_proj for the same slice chunk, so the results are
// bit-identical.  Rows of up to KW entries (images up to KW - 1 voxels wide).
// Waves per SIMD: the blocks the two staging buffers leave room for in 160 KiB
// of LDS, times NW / 4, at most 4 (the register budget of 5-12 tasks) or 6 in
// blocks of 12 waves (3-4 tasks).
constexpr int proj_rows_waves(int MODE, int ZS, int KW, int NW) {
    return (160 * 1024) / (2 * ZS * KW * (MODE == SAT_F32 ? 8 : 16)) * NW / 4 <
                   (NW == 12 ? 6 : 4)
               ? (160 * 1024) / (2 * ZS * KW * (MODE == SAT_F32 ? 8 : 16)) * NW / 4
               : (NW == 12 ? 6 : 4);
}
template <int MODE, int ZS, int T, int KW, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(proj_rows_waves(MODE, ZS, KW, NW))))
void k_proj_rows(const ProjLaunch L) {
    using E = typename Sat<MODE>::E;
    constexpr int PER = 16 / (int)sizeof(E);  // entries per lane per copy
    static_assert(KW % PER == 0, "row capacity in whole 16-byte copies");
    __shared__ __attribute__((aligned(16))) E bufA[ZS * KW];
    __shared__ __attribute__((aligned(16))) E bufB[ZS * KW];
    using LE = DDF_LDS E;
    LE* const ldsA = (LE*)bufA;
    LE* const ldsB = (LE*)bufB;
    const int lin = xcd_linear(blockIdx.x, gridDim.x);
    if (lin >= L.nwork) return;  // whole block
    int ji = 0;
    for (int t = 1; t < L.njobs; ++t)
        if (lin >= L.j[t].blk0) ji = t;
    const ProjDev& J = L.j[ji];
    const GeomDev g = J.g;
    const Subset sub = J.sub;
    const int ZT = J.zt, VB = J.vb, nug = J.nug;
    const int local = lin - J.blk0;
    const int jb = local % J.nvg;  // view block (the fastest: blocks of one chunk share rows)
    const int rest = local / J.nvg;
    const int nh = ZT / ZS;
    const int hz = rest % nh;      // which ZS slices of the chunk
    const int zc = rest / nh;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int z0 = zc * ZT;  // the chunk's slices share their angles: read at its first
    const int zs = min(z0, g.SLN - 1);

    // The block's views [jv0, jv0 + nv) and their orientation xo (block-uniform):
    // block jb of the runs of one orientation, each cut into blocks of VB views.
    // The orientation bits of 64 views at a time come from one ballot.
    int jv0 = 0, nv = 0;
    bool xo = false;
    {
        int b = jb, start = 0;
        bool ro = false, found = false;
        for (int w0 = 0; w0 < sub.count && !found; w0 += 64) {
            bool o = false;
            if (w0 + lane < sub.count) {
                const float2 cs = J.cossin[(size_t)(sub.first + (w0 + lane) * sub.stride) * g.SLN + zs];
                o = proj_xslab(cs.x, cs.y);
            }
            const unsigned long long m = __ballot(o);
            const int n = min(64, sub.count - w0);
            for (int i = 0; i < n; ++i) {
                const bool oi = (m >> i) & 1ull;
                const int jj = w0 + i;
                if (jj == 0) {
                    ro = oi;
                } else if (oi != ro) {  // the run [start, jj) ends
                    const int nb = (jj - start + VB - 1) / VB;
                    if (b < nb) {
                        jv0 = start + b * VB;
                        nv = min(VB, jj - jv0);
                        xo = ro;
                        found = true;
                        break;
                    }
                    b -= nb;
                    start = jj;
                    ro = oi;
                }
            }
        }
        if (!found) {  // the last run [start, count)
            const int nb = (sub.count - start + VB - 1) / VB;
            if (b < nb) {
                jv0 = start + b * VB;
                nv = min(VB, sub.count - jv0);
                xo = ro;
            }
        }
    }
    if (nv == 0) return;  // past the subset's blocks (whole block)

    // The wave's tasks: boundary rays of this lane; bit s of `on`: slot s holds
    // a task.
    float pa[T], pf[T], pbb[T];
    unsigned on = 0;
#pragma unroll
    for (int s = 0; s < T; ++s) {
        const int t = s * NW + wave;
        const int tv = t / nug, tug = t - tv * nug;
        pa[s] = 0.0f;
        pf[s] = 0.0f;
        pbb[s] = 0.0f;
        if (tv < nv) {
            on |= 1u << s;
            const int j = jv0 + tv;
            const float2 cs = J.cossin[(size_t)(sub.first + j * sub.stride) * g.SLN + zs];
            const int bnd = min(tug * 63 + lane, g.DNU);
            const ProjBoundary pb = proj_boundary(g.x0, g.y0, cs.x, cs.y, J.bx[bnd], J.by[bnd]);
            pa[s] = pb.a;
            pf[s] = pb.f;
            pbb[s] = pb.b;
        }
    }
    // Slots holding a task are a prefix (slot s holds task s * NW + wave, whose
    // view index grows with s): their count, wave-uniform, bounds the slot
    // loops with a scalar compare.
    const int non = __builtin_amdgcn_readfirstlane(__builtin_popcount(on));

    float acc[T][ZS];
#pragma unroll
    for (int s = 0; s < T; ++s)
#pragma unroll
        for (int q = 0; q < ZS; ++q) acc[s][q] = 0.0f;

    const int Nk = xo ? g.YN : g.XN, K = Nk + 1, Ns = xo ? g.XN : g.YN;
    const float ocS = xo ? g.ocxP : g.ocyP, ocK = xo ? g.ocyP : g.ocxP;
    const E* tab = reinterpret_cast<const E*>(xo ? J.satY : J.satX) +
                   (size_t)zc * Ns * (size_t)ZT * K + (size_t)hz * ZS * K;
    const int ncp = (K + 64 * PER - 1) / (64 * PER);  // 64-lane copies per row
    const int iB = J.ii0, iE = min(J.ii1, Ns);
    // copy c = q * ncp + cp (row q, piece cp) is issued by wave c % NW
    auto stage = [&](int ii, LE* buf) __attribute__((always_inline)) {
        const E* rows = tab + (size_t)ii * ZT * K;
#pragma unroll
        for (int q = 0; q < ZS; ++q)
            for (int cp = ((wave - q * ncp) % NW + NW) % NW; cp < ncp; cp += NW) {
                const int e0 = cp * 64 * PER;
                stage_row(rows + (size_t)q * K + e0, (DDF_LDS void*)&buf[q * KW + e0], K - e0,
                          lane);
            }
    };
    auto sum = [&](int ii, const LE* buf) __attribute__((always_inline)) {
        const float obj = slab_obj(ii, ocS, g.dx);
#pragma unroll
        for (int s = 0; s < T; ++s) {
            if (s >= non) continue;  // uniform across the wave
            const float r = slab_r(obj, pa[s], pf[s], pbb[s]);
            const Fetch f = fetch_pos<E>(slab_coord(r, g.invdx, ocK), Nk);
            const float invlen = __builtin_amdgcn_rcpf(r - lane_prev(r));
#pragma unroll
            for (int q = 0; q < ZS; ++q) {
                const Samp<MODE> Tq(lds_entry(&buf[q * KW + f.i0]), f.t);
                acc[s][q] = fmaf(Tq.minus_prev_lane(), invlen, acc[s][q]);
            }
        }
    };
    if (iB < iE) {
        stage(iB, ldsA);
        wait_vmem();
        __syncthreads();
    }
    for (int ii = iB; ii < iE; ii += 2) {
        if (ii + 1 < iE) stage(ii + 1, ldsB);
        sum(ii, ldsA);
        wait_vmem();
        __syncthreads();
        if (ii + 1 >= iE) break;
        if (ii + 2 < iE) stage(ii + 2, ldsA);
        sum(ii + 1, ldsB);
        wait_vmem();
        __syncthreads();
    }

#pragma unroll
    for (int s = 0; s < T; ++s) {
        if (s >= non) continue;
        const int t = s * NW + wave;
        const int tv = t / nug, tug = t - tv * nug;
        const int j = jv0 + tv;
        const int u = tug * 63 + lane - 1;
        if (lane < 1 || u >= g.DNU) continue;
        const float2 cs = J.cossin[(size_t)(sub.first + j * sub.stride) * g.SLN + zs];
        const float cval = proj_cval(g, cs.x, cs.y, xo, J.xds[u], J.yds[u]);
        proj_store_view<ZS>(J, g, sub, j, u, z0 + hz * ZS, acc[s], cval);
    }
}

// ------------------------------------------------------------------ index traces
__global__ __launch_bounds__(256) void k_proj_trace(const GeomDev g, const float2* __restrict__ cossin,
                                                    const float* __restrict__ bx,
                                                    const float* __restrict__ by, const Subset sub,
                                                    int32_t* __restrict__ trace) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)sub.count * g.DNU * g.SLN;
    if (idx >= n) return;
    const int z = (int)(idx % g.SLN);
    const int u = (int)((idx / g.SLN) % g.DNU);
    const int j = (int)(idx / ((long long)g.SLN * g.DNU));
    const int vg = sub.first + j * sub.stride;
    const float2 cs = cossin[(size_t)vg * g.SLN + z];
    const ProjRay ray = proj_ray(g.x0, g.y0, cs.x, cs.y, bx[u], by[u], bx[u + 1], by[u + 1]);
    const int Ns = ray.xslab ? g.XN : g.YN;
    const int Nk = ray.xslab ? g.YN : g.XN;
    const float ocS = ray.xslab ? g.ocxP : g.ocyP;
    const float ocK = ray.xslab ? g.ocyP : g.ocxP;
    const int Smax = g.XN > g.YN ? g.XN : g.YN;
    int32_t* tr = trace + (size_t)idx * Smax * 2;
    for (int ii = 0; ii < Ns; ++ii) {
        const float obj = slab_obj(ii, ocS, g.dx);
        tr[2 * ii] = texel(slab_coord(slab_r(obj, ray.aL, ray.fL, ray.bL), g.invdx, ocK), Nk);
        tr[2 * ii + 1] = texel(slab_coord(slab_r(obj, ray.aR, ray.fR, ray.bR), g.invdx, ocK), Nk);
    }
}

__global__ __launch_bounds__(256) void k_back_trace(const GeomDev g, const float2* __restrict__ src,
                                                    const Subset sub, int32_t* __restrict__ trace) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)g.XN * g.YN * g.SLN;
    if (idx >= n) return;
    const int z = (int)(idx % g.SLN);
    const int x = (int)((idx / g.SLN) % g.XN);
    const int y = (int)(idx / ((long long)g.SLN * g.XN));
    const float vx = ((float)x - g.ocxB) * g.dx;
    const float vy = ((float)y - g.ocyB) * g.dx;
    int32_t* tr = trace + (size_t)idx * sub.count * 2;
    for (int j = 0; j < sub.count; ++j) {
        const int vg = sub.first + j * sub.stride;
        const float2 sp = src[(size_t)vg * g.SLN + z];
        const BackSample bs =
            back_sample(vx, vy, sp.x, sp.y, g.inv_sid, g.invdb, g.detCntIdx, g.dx, g.backMid);
        tr[2 * j] = texel(bs.idxL, g.DNU);
        tr[2 * j + 1] = texel(bs.idxR, g.DNU);
    }
}

// ------------------------------------------------------------------ small helpers
__global__ void k_counter_inc(int* c) {
    if (threadIdx.x == 0 && blockIdx.x == 0) c[0] = c[0] + 1;
}

struct CounterLaunch {
    int n;
    int* c[kBatch];
};
__global__ void k_counter_inc_batch(const CounterLaunch L) {
    if (blockIdx.x == 0 && (int)threadIdx.x < L.n) L.c[threadIdx.x][0] = L.c[threadIdx.x][0] + 1;
}

__global__ __launch_bounds__(256) void k_fill(float* __restrict__ p, float v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ rebin
template <int RAW, int LAYOUT, int ORDER>
__global__ __launch_bounds__(256) void k_rebin(const RebinDev p, const void* __restrict__ raw,
                                               const float rawScale,
                                               const float* __restrict__ zPos,
                                               const int* __restrict__ start,
                                               const float* __restrict__ SM,
                                               float* __restrict__ Proj, float* __restrict__ Views,
                                               uint8_t* __restrict__ cut,
                                               int32_t* __restrict__ idvTrace) {
    constexpr int MAXC = 8;  // channels per thread: W <= 2048
    __shared__ int anyCut;
    // Block -> (slice s, rebinned view a).  In raw-view order the blocks that
    // read one raw view (one per slice whose window holds it) run back to back
    // on one XCD, so its frame comes from HBM once.
    const int bl = xcd_linear(blockIdx.x, gridDim.x);
    int s, a;
    if (p.lo) {
        const int k = bl / p.R;
        if (k >= p.NI) return;  // whole block
        s = p.lo[k] + (bl - k * p.R);
        if (s > p.hi[k]) return;
        a = (int)(p.imin + k - start[s]);
    } else {
        if (bl >= p.SLN * p.DefTimes) return;
        s = bl / p.DefTimes;
        a = bl - s * p.DefTimes;
    }
    if (threadIdx.x == 0) anyCut = 0;
    __syncthreads();
    const int st = start[s];
    const float z = zPos[s];
    const long long i = (long long)st + a;  // 1-based raw view of rebinned view a
    const bool valid = st >= 0 && i >= 1;
    float zz = 0.0f, angle = 0.0f;
    if (valid) {
        zz = (float)(i - 1) * p.deltaZ;
        const float lambda = (float)(i - 1) * p.delta + p.BVAngle;
        angle = mod_f(lambda, p.twopi);
    }
    // raw[(view*W + ch)*H + row]; ORDER 1: the frames in reverse view order,
    // frame f = TotalViews - view (readProj.m:43-54)
    const long long rv = ORDER == 0 ? i - 1 - p.viewOffset : p.totalViews - i;
    const size_t viewBase = valid ? (size_t)rv * p.W * p.H : 0;
    float vals[MAXC];
    int myCut = 0;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        const int jj = threadIdx.x + c * 256;
        vals[c] = 0.0f;
        if (jj < p.W) {
            int idv = 0, mk = 0;
            if (valid) {
                float v;
                const float dv = rebin_dv(z, zz, SM[jj], p.SD, p.DetCenterHreal, p.PerDetH, &v);
                idv = (int)floorf(dv);
                if (idv >= 1 && idv <= p.H - 1) {
                    const float t = dv - (float)idv;
                    const float cosA = p.SD / sqrtf(p.SD * p.SD + v * v);
                    const size_t o = viewBase + (size_t)jj * p.H + (idv - 1);
                    const size_t o1 = o + 1;  // next detector row
                    float d0, d1;
                    if constexpr (RAW == 0) {
                        const float* r = reinterpret_cast<const float*>(raw);
                        d0 = r[o];
                        d1 = r[o1];
                    } else if constexpr (RAW == 1) {  // counts -> log(scale / max(c, 1))
                        const uint16_t* r = reinterpret_cast<const uint16_t*>(raw);
                        d0 = rawScale - logf(fmaxf((float)r[o], 1.0f));
                        d1 = rawScale - logf(fmaxf((float)r[o1], 1.0f));
                    } else {  // stored line integrals times a scale (1: used unscaled)
                        const uint16_t* r = reinterpret_cast<const uint16_t*>(raw);
                        d0 = (float)r[o] * rawScale;
                        d1 = (float)r[o1] * rawScale;
                    }
                    vals[c] = cosA * (d0 * (1.0f - t) + d1 * t);
                    mk = 1;
                }
            }
            if (!mk) myCut = 1;
            if (idvTrace) idvTrace[((size_t)s * p.DefTimes + a) * p.W + jj] = idv;
        }
    }
    if (myCut) atomicOr(&anyCut, 1);
    __syncthreads();
    const bool isCut = anyCut != 0;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        const int jj = threadIdx.x + c * 256;
        if (jj < p.W) {
            Proj[((size_t)s * p.DefTimes + a) * p.W + jj] = isCut ? 0.0f : vals[c];
        }
    }
    if (threadIdx.x == 0) {
        if constexpr (LAYOUT == 0)
            Views[(size_t)s * p.DefTimes + a] = angle;
        else
            Views[(size_t)a * p.SLN + s] = angle;
        if (cut) cut[(size_t)s * p.DefTimes + a] = isCut ? 1 : 0;
    }
}

// Reference layout T[(s*DefTimes + a)*W + j] -> sinogram layout
// P[(a*W + j)*SLN + s] through 32 x 32 LDS tiles (rows of 128 bytes read and
// written).  grid (ceil(W/32), ceil(SLN/32), DefTimes), 256 threads.
__global__ __launch_bounds__(256) void k_rebin_to_sino(const float* __restrict__ T,
                                                       float* __restrict__ P, int W, int SLN,
                                                       int DefTimes) {
    __shared__ float tile[32][33];
    const int j0 = blockIdx.x * 32, s0 = blockIdx.y * 32, a = blockIdx.z;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
    for (int r = 0; r < 32; r += 8) {
        const int s = s0 + ty + r, j = j0 + tx;
        if (s < SLN && j < W) tile[ty + r][tx] = T[((size_t)s * DefTimes + a) * W + j];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 32; r += 8) {
        const int j = j0 + ty + r, s = s0 + tx;
        if (s < SLN && j < W) P[((size_t)a * W + j) * SLN + s] = tile[tx][ty + r];
    }
}

// ================================================================== launchers
#define DDF_ZT_SWITCH(ZT, ...)                                            \
    switch (ZT) {                                                         \
        case 1: { constexpr int ZTC = 1; __VA_ARGS__; } break;           \
        case 2: { constexpr int ZTC = 2; __VA_ARGS__; } break;           \
        case 4: { constexpr int ZTC = 4; __VA_ARGS__; } break;           \
        case 8: { constexpr int ZTC = 8; __VA_ARGS__; } break;           \
        case 16: { constexpr int ZTC = 16; __VA_ARGS__; } break;         \
        default: return hipErrorInvalidValue;                             \
    }

#define DDF_MODE_SWITCH(MODE, ...)                                        \
    switch (MODE) {                                                       \
        case SAT_F32: { constexpr int MC = SAT_F32; __VA_ARGS__; } break; \
        case SAT_TWOFLOAT: { constexpr int MC = SAT_TWOFLOAT; __VA_ARGS__; } break; \
        case SAT_F64: { constexpr int MC = SAT_F64; __VA_ARGS__; } break; \
        default: return hipErrorInvalidValue;                             \
    }

// Slices per SAT block: the widest G <= 4 whose LDS tiles fit in 64 KiB.  (Eight
// slices per block give 512-entry two-float rows 51 KiB of LDS, three blocks per
// CU: the two volume tables of 512x512x64 took 0.272 ms, against 0.245 ms at
// four slices per block.)
static int sat_group(int mode, int K) {
    const size_t lsz = mode == SAT_F32 ? 4 : 8;
    for (int G : {4, 2, 1})
        if ((size_t)G * K * (lsz + 4) + (size_t)256 * lsz <= 65536) return G;
    return 0;
}

// One launch over tables that share G (all rows of every table).
static hipError_t launch_sat_group(int mode, int G, SatDev* t, int n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    int maxK = 0;
    size_t bytes = 0;  // table entries the launch writes
    for (int i = 0; i < n; ++i) {
        maxK = t[i].Nk + 1 > maxK ? t[i].Nk + 1 : maxK;
        if (t[i].nrows > 0)
            bytes += (size_t)t[i].nrows * (size_t)(t[i].Nk + 1) * (size_t)(t[i].nzc * t[i].ZT) *
                     (mode == SAT_F32 ? 8 : 16);
    }
    const bool nt = bytes >= kSatNtBytes;
    SatLaunch L{};
    long long nb = 0;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (t[i].nrows <= 0) continue;
        t[i].nzg = (t[i].nzc * t[i].ZT + G - 1) / G;
        t[i].blk0 = (int)nb;
        nb += (long long)t[i].nrows * t[i].nzg;
        L.j[m++] = t[i];
    }
    L.njobs = m;
    if (nb == 0) return hipSuccess;
    if (nb > 0x7fffff00LL) return hipErrorInvalidValue;
    L.nwork = (int)nb;
    if ((maxK - 1 + 256 / G - 1) / (256 / G) > kSatSeg) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)xcd_grid((int)nb);
    const size_t lsz = mode == SAT_F32 ? 4 : 8;
    const size_t lds = (size_t)G * maxK * (lsz + 4) + (size_t)256 * lsz;
#define DDF_SATV(GG)                                                                   \
    if (nt) {                                                                          \
        DDF_MODE_SWITCH(mode, (k_sat_rows<MC, GG, true><<<grid, 256, lds, st>>>(L)));  \
    } else {                                                                           \
        DDF_MODE_SWITCH(mode, (k_sat_rows<MC, GG, false><<<grid, 256, lds, st>>>(L))); \
    }
    switch (G) {
        case 4: DDF_SATV(4); break;
        case 2: DDF_SATV(2); break;
        default: DDF_SATV(1); break;
    }
#undef DDF_SATV
    return hipGetLastError();
}

// One launch of k_sat_tiles over tables that share its slice group G.
static hipError_t launch_sat_tile_group(int mode, int G, SatDev* t, int n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    SatLaunch L{};
    long long nb = 0;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (t[i].nrows <= 0) continue;
        t[i].nzg = (t[i].nzc * t[i].ZT + G - 1) / G;
        t[i].blk0 = (int)nb;
        nb += (long long)t[i].nrows * t[i].nzg;
        L.j[m++] = t[i];
    }
    L.njobs = m;
    size_t bytes = 0;  // table entries the launch writes
    for (int i = 0; i < m; ++i)
        bytes += (size_t)L.j[i].nrows * (size_t)(L.j[i].Nk + 1) * (size_t)(L.j[i].nzc * L.j[i].ZT) *
                 (mode == SAT_F32 ? 8 : 16);
    L.nt = bytes >= kSatTileNtBytes;
    if (nb == 0) return hipSuccess;
    if (nb > 0x7fffff00LL) return hipErrorInvalidValue;
    L.nwork = (int)nb;
    const unsigned grid = (unsigned)xcd_grid((int)nb);
    bool v4 = true;  // every table's volume: slices a multiple of 4, 16-byte aligned
    for (int i = 0; i < m; ++i)
        v4 = v4 && L.j[i].SLN % 4 == 0 && reinterpret_cast<uintptr_t>(L.j[i].in) % 16 == 0;
    if (G == 32 && v4) {
        DDF_MODE_SWITCH(mode, (k_sat_tiles<MC, 32, true><<<grid, 256, 0, st>>>(L)));
    } else if (G == 32) {
        DDF_MODE_SWITCH(mode, (k_sat_tiles<MC, 32, false><<<grid, 256, 0, st>>>(L)));
    } else if (v4) {
        DDF_MODE_SWITCH(mode, (k_sat_tiles<MC, 16, true><<<grid, 256, 0, st>>>(L)));
    } else {
        DDF_MODE_SWITCH(mode, (k_sat_tiles<MC, 16, false><<<grid, 256, 0, st>>>(L)));
    }
    return hipGetLastError();
}

// Slices per block of the tiled builder for a volume table of SLN slices: 16
// or 32, or 0 for the row-per-block builder (tables of fewer than 16 slices,
// and sinogram tables).  The two builders group a row's sums differently, so
// the choice follows the volume's own slice count, never its padded count
// nzc * ZT, which depends on the slice chunk (13 slices in chunks of 2 or of
// 4 would otherwise get different tables).  The tiled builder's sums do not
// depend on its slices per block, so that choice is free: 16 for 16-byte
// entries (two-float, f64), whose rows then leave as 1 KiB runs; for f32
// entries 32 unless 16 pads the slice count less.  Measured on MI355X
// (both volume tables per call, 32 -> 16 slices per block): 512x512x64
// two-float 0.150 -> 0.144 ms, f32 0.097 -> 0.107 ms; 512x512x200 in one-slice
// chunks two-float 0.593 -> 0.535 ms, f32 0.375 -> 0.345 ms.  Row per block
// vs tiled at 512x512x64, 2304 views (two-float): volume tables 0.236 vs
// 0.157 ms; the sinogram table 0.534 vs 0.601 ms (both builders then move
// 3.6-4.3 TB/s, most of it written entries).
static int sat_tile_group(int mode, const SatDev& t) {
    if (!t.tiled || t.SLN < 16) return 0;
    if (mode == SAT_F32 && t.SLN >= 32 && (t.SLN + 31) / 32 * 32 <= (t.SLN + 15) / 16 * 16) return 32;
    return 16;
}

// Up to 2*kBatch tables.  The builder and its slices per block fix how a row
// is split into scan segments, and so the summation order: they are chosen
// from each table's own shape, never from the other tables of the launch, so
// a table's entries are the same whatever it is batched with.
static hipError_t launch_sat_tables(int mode, SatDev* t, int n, hipStream_t st) {
    for (int G : {32, 16}) {
        SatDev grp[2 * kBatch];
        int m = 0;
        for (int i = 0; i < n; ++i)
            if (sat_tile_group(mode, t[i]) == G) grp[m++] = t[i];
        const hipError_t e = launch_sat_tile_group(mode, G, grp, m, st);
        if (e != hipSuccess) return e;
    }
    for (int G : {4, 2, 1}) {
        SatDev grp[2 * kBatch];
        int m = 0;
        for (int i = 0; i < n; ++i)
            if (sat_tile_group(mode, t[i]) == 0 && sat_group(mode, t[i].Nk + 1) == G) grp[m++] = t[i];
        const hipError_t e = launch_sat_group(mode, G, grp, m, st);
        if (e != hipSuccess) return e;
    }
    for (int i = 0; i < n; ++i)
        if (sat_tile_group(mode, t[i]) == 0 && sat_group(mode, t[i].Nk + 1) == 0)
            return hipErrorInvalidValue;
    return hipSuccess;
}

hipError_t launch_sat_vol_batch(int mode, const SatVolJob* jobs, int n, hipStream_t st) {
    for (int i0 = 0; i0 < n; i0 += kBatch) {
        SatDev t[2 * kBatch];
        int m = 0;
        for (int i = i0; i < n && i < i0 + kBatch; ++i) {
            const SatVolJob& J = jobs[i];
            // rows y, prefix along x; rows x, prefix along y
            t[m++] = SatDev{J.vol, J.mask, J.satX, J.g.YN, J.g.XN, J.g.XN, 1, J.g.SLN, J.ZT, J.nzc, 0, 0, 1};
            t[m++] = SatDev{J.vol, J.mask, J.satY, J.g.XN, J.g.YN, 1, J.g.XN, J.g.SLN, J.ZT, J.nzc, 0, 0, 1};
        }
        const hipError_t e = launch_sat_tables(mode, t, m, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sat_sino_batch(int mode, const SatSinoJob* jobs, int n, hipStream_t st) {
    for (int i0 = 0; i0 < n; i0 += 2 * kBatch) {
        SatDev t[2 * kBatch];
        int m = 0;
        for (int i = i0; i < n && i < i0 + 2 * kBatch; ++i) {
            const SatSinoJob& J = jobs[i];  // rows = views, prefix along u
            t[m++] = SatDev{J.r, nullptr, J.P, J.PNs, J.DNU, J.DNU, 1, J.SLN, J.ZT, J.nzc, 0, 0, 0};
        }
        const hipError_t e = launch_sat_tables(mode, t, m, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_sat_vol(int mode, int ZT, const GeomDev& g, int nzc, const float* vol,
                          const uint8_t* mask, void* satX, void* satY, hipStream_t st) {
    const SatVolJob j{g, nzc, ZT, vol, mask, satX, satY};
    return launch_sat_vol_batch(mode, &j, 1, st);
}

hipError_t launch_sat_sino(int mode, int ZT, int DNU, int SLN, int nzc, int PNs, const float* r,
                           void* P, hipStream_t st) {
    if (PNs == 0) return hipSuccess;
    const SatSinoJob j{DNU, SLN, nzc, ZT, PNs, r, P};
    return launch_sat_sino_batch(mode, &j, 1, st);
}

// Shapes of k_proj_rows by the plan's tasks-per-wave value: 9 or 12 tasks in
// blocks of 4 waves, 5 or 6 in blocks of 8, 3 or 4 in blocks of 12.
static int proj_rows_nw(int V) { return V >= 9 ? 4 : V >= 5 ? 8 : 12; }
static int proj_rows_t(int V) { return V; }

// Blocks per slice chunk of k_proj_rows: the subset's runs of views of one slab
// orientation, each cut into blocks of VB views (the most over the chunks;
// without the host angle table, one block per view).
static int proj_rows_nvg(const ProjJob& J, int VB) {
    const int n = J.sub.count;
    if (!J.hcs) return n;
    int most = 0;
    for (int zc = 0; zc < J.nzc; ++zc) {
        const int zs = std::min(zc * J.ZT, J.g.SLN - 1);
        int blocks = 0, start = 0;
        bool ro = false;
        for (int j = 0; j <= n; ++j) {
            bool o = false;
            if (j < n) {
                const size_t v = (size_t)(J.sub.first + (long long)j * J.sub.stride) * J.g.SLN + zs;
                o = proj_xslab(J.hcs[2 * v], J.hcs[2 * v + 1]);
            }
            if (j == 0) {
                ro = o;
            } else if (j == n || o != ro) {
                blocks += (j - start + VB - 1) / VB;
                start = j;
                ro = o;
            }
        }
        most = std::max(most, blocks);
    }
    return most;
}

// Row capacity of k_proj_rows for rows of K entries (0: too long).
static int proj_rows_kw(int mode, int K) {
    const int per = mode == SAT_F32 ? 2 : 1;
    const int k = (K + per - 1) / per * per;
    return k <= 520 ? 520 : k <= 776 ? 776 : 0;
}

// Launch shape checks of the projector kernels: stage 0/1 take 1, 2 or 4 views
// per wave; the whole-row kernel (stage 2: 4 slices per block, 3: 2) takes 3,
// 4, 5, 6, 9 or 12 tasks per wave, a slice chunk that is a multiple of its slices,
// rows within its capacity and a view block of at least one view.
bool proj_shape_ok(int mode, const GeomDev& g, int ZT, int V, int stage) {
    if (stage == 0 || stage == 1) return V == 1 || V == 2 || V == 4;
    if (stage != 2 && stage != 3) return false;
    const int ZS = stage == 2 ? 4 : 2;
    const int nug = (g.DNU + 62) / 63;
    return (V == 3 || V == 4 || V == 5 || V == 6 || V == 9 || V == 12) && ZT >= ZS &&
           ZT % ZS == 0 &&
           proj_rows_kw(mode, (g.XN > g.YN ? g.XN : g.YN) + 1) > 0 &&
           proj_rows_nw(V) * proj_rows_t(V) >= nug;
}

// Projector launches: jobs of one slice chunk ZT go together, kBatch at a time;
// pass p of every job that still has slabs left shares one launch.
static hipError_t launch_proj_group(int mode, int ZT, int V, int stage,
                                    const ProjJob* const* jobs, int n, hipStream_t st) {
    int maxPass = 0;
    int T[kBatch], Nmax[kBatch];
    int KW = 0;  // row capacity of the whole-row kernel: the launch's longest rows
    for (int i = 0; i < n; ++i) {
        Nmax[i] = jobs[i]->g.XN > jobs[i]->g.YN ? jobs[i]->g.XN : jobs[i]->g.YN;
        T[i] = (jobs[i]->slabTile <= 0 || jobs[i]->slabTile > Nmax[i]) ? Nmax[i] : jobs[i]->slabTile;
        const int np = (Nmax[i] + T[i] - 1) / T[i];
        maxPass = np > maxPass ? np : maxPass;
        if (!proj_shape_ok(mode, jobs[i]->g, ZT, V, stage)) return hipErrorInvalidValue;
        const int kw = proj_rows_kw(mode, Nmax[i] + 1);
        KW = kw > KW ? kw : KW;
    }
    const bool rows = stage >= 2;
    const int ZS = stage == 2 ? 4 : 2;
    for (int p = 0; p < maxPass; ++p) {
        ProjLaunch L{};
        long long nw = 0;
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const ProjJob& J = *jobs[i];
            const int ii0 = p * T[i];
            if (ii0 >= Nmax[i] || J.sub.count <= 0) continue;
            ProjDev& d = L.j[m++];
            d.g = J.g;
            d.sub = J.sub;
            d.nzc = J.nzc;
            d.nug = (J.g.DNU + 62) / 63;  // 63 detectors per wave (64 boundaries)
            d.zt = ZT;
            d.vb = rows ? proj_rows_nw(V) * proj_rows_t(V) / d.nug : V;
            d.nvg = rows ? proj_rows_nvg(J, d.vb) : (J.sub.count + d.vb - 1) / d.vb;
            if (stage == 1) d.nvg = (d.nvg + 3) / 4;  // blocks of four view groups
            d.blk0 = (int)nw;
            d.ii0 = ii0;
            d.ii1 = ii0 + T[i];
            d.first = p == 0;
            d.last = ii0 + T[i] >= Nmax[i];
            d.epi = J.epi;
            d.cossin = J.cossin;
            d.bx = J.bx;
            d.by = J.by;
            d.xds = J.xds;
            d.yds = J.yds;
            d.satX = J.satX;
            d.satY = J.satY;
            d.out = J.out;
            d.meas = J.meas;
            d.rowSum = J.rowSum;
            nw += rows ? (long long)d.nvg * d.nzc * (ZT / ZS) : (long long)d.nug * d.nvg * d.nzc;
        }
        if (m == 0) continue;
        if (nw > 0x7fffff00LL) return hipErrorInvalidValue;
        L.njobs = m;
        L.nwork = (int)nw;
        dim3 grid((unsigned)xcd_grid(L.nwork));
#define DDF_PROJ(VV)                                                                       \
    if (stage) {                                                                           \
        DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, (k_proj_lds<MC, ZTC, VV><<<grid, 256, 0, st>>>(L)))) \
    } else {                                                                               \
        DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, (k_proj<MC, ZTC, VV><<<grid, 64, 0, st>>>(L))))      \
    }
#define DDF_ROWS(ZSC, TT, NWC)                                                             \
    if (KW == 520) {                                                                       \
        DDF_MODE_SWITCH(mode, (k_proj_rows<MC, ZSC, TT, 520, NWC><<<grid, 64 * NWC, 0, st>>>(L))) \
    } else {                                                                               \
        DDF_MODE_SWITCH(mode, (k_proj_rows<MC, ZSC, TT, 776, NWC><<<grid, 64 * NWC, 0, st>>>(L))) \
    }
#define DDF_ROWS_T(ZSC)                                                                    \
    if (V == 12) {                                                                         \
        DDF_ROWS(ZSC, 12, 4)                                                               \
    } else if (V == 9) {                                                                   \
        DDF_ROWS(ZSC, 9, 4)                                                                \
    } else if (V == 6) {                                                                   \
        DDF_ROWS(ZSC, 6, 8)                                                                \
    } else if (V == 5) {                                                                   \
        DDF_ROWS(ZSC, 5, 8)                                                                \
    } else if (V == 4) {                                                                   \
        DDF_ROWS(ZSC, 4, 12)                                                               \
    } else {                                                                               \
        DDF_ROWS(ZSC, 3, 12)                                                               \
    }
        if (rows) {
            if (ZS == 4) {
                DDF_ROWS_T(4);
            } else {
                DDF_ROWS_T(2);
            }
        } else if (V == 4) {
            DDF_PROJ(4);
        } else if (V == 2) {
            DDF_PROJ(2);
        } else {
            DDF_PROJ(1);
        }
#undef DDF_ROWS_T
#undef DDF_ROWS
#undef DDF_PROJ
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Batched launches stop adding jobs once the launch holds fillWaves projector
// waves (fillBlocks backprojector blocks): small reconstructions share
// launches, large ones keep the GPU (and its caches) to themselves.  Grouping
// never changes a job's arithmetic.
hipError_t launch_proj_batch(int mode, const ProjJob* jobs, int n, hipStream_t st,
                             long long fillWaves) {
    for (int i = 0; i < n; ++i)
        if (jobs[i].sub.count > 0 && !proj_shape_ok(mode, jobs[i].g, jobs[i].ZT, jobs[i].V,
                                                    jobs[i].stage))
            return hipErrorInvalidValue;
    for (int stage : {0, 1, 2, 3})
        for (int V : {1, 2, 3, 4, 5, 6, 9, 12}) {
            for (int ZT : {1, 2, 4, 8, 16}) {
                const ProjJob* grp[kBatch];
                int m = 0;
                long long work = 0;
                for (int i = 0; i < n; ++i) {
                    const ProjJob& J = jobs[i];
                    if (J.ZT != ZT || J.V != V || J.stage != stage || J.sub.count <= 0) continue;
                    const long long w =
                        (long long)((J.g.DNU + 62) / 63) * ((J.sub.count + V - 1) / V) * J.nzc;
                    if (m == kBatch || (m > 0 && work + w > fillWaves)) {
                        const hipError_t e = launch_proj_group(mode, ZT, V, stage, grp, m, st);
                        if (e != hipSuccess) return e;
                        m = 0;
                        work = 0;
                    }
                    grp[m++] = &J;
                    work += w;
                }
                if (m) {
                    const hipError_t e = launch_proj_group(mode, ZT, V, stage, grp, m, st);
                    if (e != hipSuccess) return e;
                }
            }
        }
    return hipSuccess;
}

hipError_t launch_proj(int mode, int ZT, int epi, const GeomDev& g, int nzc, const float2* cossin,
                       const float* bx, const float* by, const float* xds, const float* yds,
                       const void* satX, const void* satY, Subset sub, float* out,
                       const float* meas, const float* rowSum, int slabTile, int viewsPerWave,
                       int stage, hipStream_t st, const float* hcs) {
    if (sub.count <= 0) return hipSuccess;
    const ProjJob j{g,   sub,  nzc, ZT,   slabTile, epi,  viewsPerWave, cossin, bx,    by,
                    xds, yds,  satX, satY, out,     meas, rowSum,       stage,  hcs};
    return launch_proj_batch(mode, &j, 1, st);
}

// Views per staged group of k_back_lds: 1, 2 or 4, with the two LDS buffers of
// 2 * NV * ZT KiB within 64 KiB.
static bool back_stage_ok(int ZT, int NV) {
    return NV == 0 || ((NV == 1 || NV == 2 || NV == 4) && NV * ZT <= 32);
}

template <int MC, int ZTC>
static hipError_t launch_back_kernel(int NV, dim3 grid, const BackLaunch& L, hipStream_t st) {
    switch (NV) {
        case 0: k_back<MC, ZTC><<<grid, 256, 0, st>>>(L); break;
        case 1: k_back_lds<MC, ZTC, 1><<<grid, 256, 0, st>>>(L); break;
        case 2:
            if constexpr (2 * ZTC <= 32) k_back_lds<MC, ZTC, 2><<<grid, 256, 0, st>>>(L);
            else return hipErrorInvalidValue;
            break;
        case 4:
            if constexpr (4 * ZTC <= 32) k_back_lds<MC, ZTC, 4><<<grid, 256, 0, st>>>(L);
            else return hipErrorInvalidValue;
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static hipError_t launch_back_group(int mode, int ZT, int NV, const BackJob* const* jobs, int n,
                                    hipStream_t st) {
    BackLaunch L{};
    long long nw = 0;
    int m = 0;
    for (int i = 0; i < n; ++i) {
        const BackJob& J = *jobs[i];
        BackDev& d = L.j[m++];
        d.g = J.g;
        d.sub = J.sub;
        d.nzc = J.nzc;
        d.ntx = (J.g.XN + 15) / 16;
        d.nty = (J.g.YN + 15) / 16;
        d.blk0 = (int)nw;
        d.epi = J.epi;
        d.tabViews = J.tabViews > 0 ? J.tabViews : J.sub.count;
        d.accum = J.accum;
        d.src = J.src;
        d.mask = J.mask;
        d.P = J.P;
        d.vol = J.vol;
        d.colSum = J.colSum;
        d.xprev = J.xprev;
        d.momTab = J.momTab;
        d.counter = J.counter;
        nw += (long long)d.ntx * d.nty * d.nzc;
    }
    if (nw == 0) return hipSuccess;
    if (nw > 0x7fffff00LL) return hipErrorInvalidValue;
    L.njobs = m;
    L.nwork = (int)nw;
    dim3 grid((unsigned)xcd_grid(L.nwork));
    hipError_t e = hipErrorInvalidValue;
    DDF_MODE_SWITCH(mode, DDF_ZT_SWITCH(ZT, (e = launch_back_kernel<MC, ZTC>(NV, grid, L, st))));
    return e;
}

hipError_t launch_back_batch(int mode, const BackJob* jobs, int n, hipStream_t st,
                             long long fillBlocks) {
    for (int i = 0; i < n; ++i)
        if (!back_stage_ok(jobs[i].ZT, jobs[i].NV)) return hipErrorInvalidValue;
    for (int ZT : {1, 2, 4, 8, 16}) {
        for (int NV : {0, 1, 2, 4}) {
            const BackJob* grp[kBatch];
            int m = 0;
            long long work = 0;
            for (int i = 0; i < n; ++i) {
                const BackJob& J = jobs[i];
                if (J.ZT != ZT || J.NV != NV) continue;
                const long long w =
                    (long long)((J.g.XN + 15) / 16) * ((J.g.YN + 15) / 16) * J.nzc;
                if (m == kBatch || (m > 0 && work + w > fillBlocks)) {
                    const hipError_t e = launch_back_group(mode, ZT, NV, grp, m, st);
                    if (e != hipSuccess) return e;
                    m = 0;
                    work = 0;
                }
                grp[m++] = &J;
                work += w;
            }
            if (m) {
                const hipError_t e = launch_back_group(mode, ZT, NV, grp, m, st);
                if (e != hipSuccess) return e;
            }
        }
    }
    return hipSuccess;
}

hipError_t launch_back(int mode, int ZT, int epi, const GeomDev& g, int nzc, const float2* src,
                       const uint8_t* mask, const void* P, Subset sub, float* vol,
                       const float* colSum, float* xprev, const float* momTab,
                       const int* iterCounter, int stageViews, hipStream_t st, int viewPass) {
    if (viewPass <= 0 || viewPass >= sub.count || epi != BACK_PLAIN) {
        const BackJob j{g,      sub,   nzc,    ZT,          epi,       src, mask, P,
                        vol,    colSum, xprev, momTab,      iterCounter, stageViews};
        return launch_back_batch(mode, &j, 1, st);
    }
    // View passes: pass p sums views [p*viewPass, (p+1)*viewPass) of the
    // subset on top of the sums the previous passes left in vol (the same
    // additions in the same order as one pass, so the same results).
    const size_t eb = mode == SAT_F32 ? 8 : 16;
    for (int j0 = 0; j0 < sub.count; j0 += viewPass) {
        const Subset sp{sub.first + j0 * sub.stride, sub.stride, std::min(viewPass, sub.count - j0)};
        const void* Pp = static_cast<const char*>(P) + (size_t)j0 * ZT * (g.DNU + 1) * eb;
        BackJob j{g,   sp,     nzc,   ZT,     epi,         src,        mask, Pp,
                  vol, colSum, xprev, momTab, iterCounter, stageViews};
        j.tabViews = sub.count;
        j.accum = j0 > 0;
        const hipError_t e = launch_back_batch(mode, &j, 1, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_counter_inc_batch(int* const* counters, int n, hipStream_t st) {
    for (int i0 = 0; i0 < n; i0 += kBatch) {
        CounterLaunch L{};
        for (int i = i0; i < n && i < i0 + kBatch; ++i) L.c[L.n++] = counters[i];
        k_counter_inc_batch<<<1, 64, 0, st>>>(L);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_proj_trace(const GeomDev& g, const float2* cossin, const float* bx,
                             const float* by, Subset sub, int32_t* trace, hipStream_t st) {
    const long long n = (long long)sub.count * g.DNU * g.SLN;
    if (n == 0) return hipSuccess;
    k_proj_trace<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, cossin, bx, by, sub, trace);
    return hipGetLastError();
}

hipError_t launch_back_trace(const GeomDev& g, const float2* src, Subset sub, int32_t* trace,
                             hipStream_t st) {
    const long long n = (long long)g.XN * g.YN * g.SLN;
    if (n == 0) return hipSuccess;
    k_back_trace<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g, src, sub, trace);
    return hipGetLastError();
}

hipError_t launch_counter_inc(int* counter, hipStream_t st) {
    k_counter_inc<<<1, 64, 0, st>>>(counter);
    return hipGetLastError();
}

hipError_t launch_fill(float* p, float v, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_rebin(const RebinDev& p, int rawKind, int rawOrder, int outLayout,
                        const void* raw, float countScale, const float* zPos, const int* start,
                        const float* SM, float* Proj, float* Views, uint8_t* cut,
                        int32_t* idvTrace, float* tmp, hipStream_t st) {
    if (p.W > 8 * 256) return hipErrorInvalidValue;
    const long long nb = p.lo ? (long long)p.NI * p.R : (long long)p.SLN * p.DefTimes;
    if (nb == 0) return hipSuccess;
    if (nb > 0x7fffff00LL || (long long)p.SLN * p.DefTimes > 0x7fffff00LL)
        return hipErrorInvalidValue;
    if (rawOrder == 1 && p.viewOffset != 0) return hipErrorInvalidValue;
    if (outLayout == 1 && !tmp) return hipErrorInvalidValue;
    const float scale = rawKind == 1 ? logf(countScale) : countScale;
    const unsigned grid = (unsigned)xcd_grid((int)nb);
    // the gather writes the reference layout (rows of W contiguous samples);
    // the sinogram layout is a transpose of it
    float* const gProj = outLayout == 1 ? tmp : Proj;
#define DDF_REBIN(RK, LY, OR)                                                                   \
    k_rebin<RK, LY, OR><<<grid, 256, 0, st>>>(p, raw, scale, zPos, start, SM, gProj, Views, cut, \
                                              idvTrace)
#define DDF_REBIN_O(RK, LY)                  \
    do {                                     \
        if (rawOrder == 0)                   \
            DDF_REBIN(RK, LY, 0);            \
        else if (rawOrder == 1)              \
            DDF_REBIN(RK, LY, 1);            \
        else                                 \
            return hipErrorInvalidValue;     \
    } while (0)
    if (outLayout != 0 && outLayout != 1) return hipErrorInvalidValue;
    // the samples go out in the reference layout either way; LY selects the
    // layout of the view angles
    if (rawKind == 0) {
        if (outLayout == 0) DDF_REBIN_O(0, 0); else DDF_REBIN_O(0, 1);
    } else if (rawKind == 1) {
        if (outLayout == 0) DDF_REBIN_O(1, 0); else DDF_REBIN_O(1, 1);
    } else if (rawKind == 2) {
        if (outLayout == 0) DDF_REBIN_O(2, 0); else DDF_REBIN_O(2, 1);
    } else {
        return hipErrorInvalidValue;
    }
#undef DDF_REBIN_O
#undef DDF_REBIN
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || outLayout == 0) return e;
    const dim3 tg((unsigned)((p.W + 31) / 32), (unsigned)((p.SLN + 31) / 32), (unsigned)p.DefTimes);
    k_rebin_to_sino<<<tg, 256, 0, st>>>(tmp, Proj, p.W, p.SLN, p.DefTimes);
    return hipGetLastError();
}

}  // namespace ddf
