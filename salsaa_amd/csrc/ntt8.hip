// ntt8.hip -- batched NTT kernels for d = 64 with 8 threads per ring element
// (see ntt8.hpp), a block of 1024 threads staging 128 elements through LDS
// with coalesced 16-byte global accesses.
//
//   k_ntt_inc<false/true>  Coefficient <-> IncompleteNTT in place
//                          (cyclotomic_ring.rs:483-511, 527-549)
//   k_ntt_full<false/true> Coefficient <-> NTT in place (cyclotomic_ring.rs:513-525)
//   k_ntt_full_sm          Coefficient rows -> full NTT written slot-major as
//                          exact doubles [64][rows][m] (the commit's W layout)
//   k_decomp_sm            decompose_matrix_by_chunks fused with the limbs'
//                          full NTT: writes the limb witness (Coefficient) and
//                          the limbs' slot-major NTT in one pass over W
#include <algorithm>

#include "ntt8.hpp"

namespace salsaa {

namespace {

constexpr int EL = 128;      // ring elements per block
constexpr int NT = 1024;     // threads per block (8 per element)
constexpr int SROW = 72;     // LDS row stride in u64 (64 + 8 pad)


// LDS row layouts (row stride SROW u64): "plain" keeps entry k at k; "skew"
// puts entry k at k + k/8, which makes the 8-consecutive-entry accesses of
// phase B and the stride-4/8 accesses of phase A bank-conflict free.
__device__ __forceinline__ int skew(int k) { return k + (k >> 3); }

template <bool SKEW>
__device__ __forceinline__ void load_block(uint64_t *lds, const uint64_t *__restrict__ src, int nel,
                                           int tid) {
    const ulonglong2 *s2 = reinterpret_cast<const ulonglong2 *>(src);
    for (int c = tid; c < nel * 32; c += NT) {
        const int e = c >> 5, k = (c & 31) * 2;
        const ulonglong2 v = s2[c];
        if (SKEW) {
            lds[e * SROW + skew(k)] = v.x;
            lds[e * SROW + skew(k + 1)] = v.y;
        } else {
            *reinterpret_cast<ulonglong2 *>(&lds[e * SROW + k]) = v;
        }
    }
}

template <bool SKEW>
__device__ __forceinline__ void store_block(uint64_t *__restrict__ dst, const uint64_t *lds, int nel,
                                            int tid) {
    ulonglong2 *d2 = reinterpret_cast<ulonglong2 *>(dst);
    for (int c = tid; c < nel * 32; c += NT) {
        const int e = c >> 5, k = (c & 31) * 2;
        ulonglong2 v;
        if (SKEW) {
            v.x = lds[e * SROW + skew(k)];
            v.y = lds[e * SROW + skew(k + 1)];
        } else {
            v = *reinterpret_cast<const ulonglong2 *>(&lds[e * SROW + k]);
        }
        d2[c] = v;
    }
}

// Slot-major staging of a block's 64 x 128 transformed values: slot s,
// element e sits at wpos(s, e), i.e. row s with each run of 32 elements
// rotated by r(s, e / 32) = 8 ((s / 8) % 4) + 2 (s % 8) + (e / 32) % 2.  The
// rotation spreads the three access patterns over the LDS banks: the NTT's
// writes (8 consecutive e by slots 8 apart), the plane packing's reads
// (8 consecutive e per thread, 4 slots x 16 runs of 8 per wave) and the
// slot-major store's reads (64 consecutive e of one slot).
template <int L = EL>
__device__ __forceinline__ int wpos(int s, int e) {
    const int r = 8 * ((s >> 3) & 3) + 2 * (s & 7) + ((e >> 5) & 1);
    return s * L + (e & ~31) + ((e + r) & 31);
}

// The same staging for the per-wave NTT pass (128-element rows), as an XOR
// swizzle of each run of 32 elements.  gfx950 serves an 8-byte LDS access
// conflict-free when each half of the wave hits 32 distinct u64 modulo 32
// (measured); this layout does that for both of the pass's patterns: slot
// writes (8 elements of a wave x slots 8 h + x, h < 8) and the packers'
// reads (2 slots x 16 runs of 8 elements per half wave).
__device__ __forceinline__ int wpos_w(int s, int e) {
    const int f = ((e >> 5) & 3) ^ (4 * ((s ^ (s >> 5)) & 1)) ^ (8 * ((s >> 3) & 3));
    return s * EL + (e & ~31) + ((e & 31) ^ f);
}

// Balanced base-256 digits of a residue x < 2^50: x = sum_u d_u 256^u with
// d_u in [-128, 127] (7 digits).  With y = x + sum_u 128 * 256^u the bytes
// b_u of y give d_u = b_u - 128, whose int8 bit pattern is b_u XOR 0x80.
constexpr uint64_t DIGIT_BIAS = 0x0080808080808080ULL;

// Fragment-tiled digit planes for the int8 MFMA commit.  A matrix of R value
// rows has 7R digit rows (row r*7 + u holds digit u), grouped in tiles of 16;
// columns are grouped in K-tiles of 64.  Fragment (s, kt, rt) is the 1 KiB
// block at ((s * KT + kt) * RT + rt) * 1024, RT = ceil(7R / 16), KT = m
// rounded up to 128, over 64, stored row-major (digit row dr % 16 at
// (dr % 16) * 64).  So one (slot, K-tile) step of the whole matrix is a
// row-major [16 RT][64] byte block: the row tiles a commit block needs for
// one K step are consecutive, and the 7 digit rows of one value row are 448
// contiguous bytes, which this writer fills with coalesced 8-byte stores.
// Columns in [m, 64 KT) are written as zero; the padding rows of the last row
// tile are never written and only feed products that are discarded.
// Each thread packs one digit of 8 columns into one 8-byte store.
// `base` is the fragment byte of (slot 0, K tile i0 / 64, value row row), i.e.
// dst + ((i0 / 64) RT) 1024 + 448 row; slots are sstride = KT RT 1024 bytes
// apart and K tiles kstride = RT 1024.
// Item c of L / 8 runs of 8 columns per slot (L = the tile's columns); WL:
// the staging is in the wpos_w layout.
template <int L = EL, bool WL = false>
__device__ __forceinline__ void store_planes_one(int8_t *__restrict__ base, size_t sstride,
                                                 size_t kstride, const uint64_t *work, int nel,
                                                 int c) {
    {
        const int s = c / (L / 8), c8 = c % (L / 8);
        // two halves of 4 columns each, so that only 4 values are live at once
        uint32_t p[7][2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int e = 8 * c8 + 4 * h + q;
                uint64_t y = 0;  // zero padding: every digit 0
                if (e < nel)
                    y = (d2u(bitsd(work[WL ? wpos_w(s, e) : wpos<L>(s, e)])) + DIGIT_BIAS) ^
                        0x8080808080808080ULL;
                lo[q] = (uint32_t)y;
                hi[q] = (uint32_t)(y >> 32);
            }
            uint32_t dg[7];
            bytes_transpose4(lo, hi, dg);
#pragma unroll
            for (int u = 0; u < 7; u++) p[u][h] = dg[u];
        }
        // nontemporal: the planes are read back by the commit long after,
        // from HBM, so they need not take up cache on the way out
        int8_t *fbase = base + s * sstride + (c8 >> 3) * kstride + (c8 & 7) * 8;
#pragma unroll
        for (int u = 0; u < 7; u++)
            __builtin_nontemporal_store(
                (unsigned long long)p[u][0] | ((unsigned long long)p[u][1] << 32),
                reinterpret_cast<unsigned long long *>(fbase + u * 64));
    }
}
__device__ __forceinline__ void store_planes(int8_t *__restrict__ base, size_t sstride,
                                             size_t kstride, const uint64_t *work, int nel,
                                             int tid) {
    for (int c = tid; c < 64 * (EL / 8); c += NT)
        store_planes_one(base, sstride, kstride, work, nel, c);
}

__device__ __forceinline__ size_t plane_kstride(size_t rows) { return (rows * 7 + 15) / 16 * 1024; }
__device__ __forceinline__ size_t plane_sstride(size_t rows, size_t ktiles) {
    return ktiles * plane_kstride(rows);
}
__device__ __forceinline__ int8_t *plane_base(int8_t *dst, size_t rows, size_t row, size_t i0) {
    return dst + (i0 / 64) * plane_kstride(rows) + row * 7 * 64;
}

// Tile staging for the persistent NTT: thread tid moves chunks tid + NT i
// (i < 4) of a tile of EL ring elements, held in registers so the next tile's
// loads are in flight while this one is computed.  Branch-free, so the loads
// stay in flight: chunks past the tile's nel elements re-read chunk 0 and land
// in LDS rows of elements that are neither transformed nor stored.
__device__ __forceinline__ ulonglong2 tile_chunk(const uint64_t *__restrict__ src, int c, int nel) {
    return reinterpret_cast<const ulonglong2 *>(src)[c < nel * 32 ? c : 0];
}

template <bool SKEW>
__device__ __forceinline__ void tile_put1(uint64_t *lds, ulonglong2 v, int c) {
    const int e = c >> 5, k = (c & 31) * 2;
    if (SKEW) {
        lds[e * SROW + skew(k)] = v.x;
        lds[e * SROW + skew(k + 1)] = v.y;
    } else {
        *reinterpret_cast<ulonglong2 *>(&lds[e * SROW + k]) = v;
    }
}

// IncompleteNTT: thread h of element e works on half p = h >> 2 as t = h & 3.
// Half p, position i lives at 36 p + skew(i) in the skewed half layout.
// Persistent blocks: a block transforms tiles blockIdx.x, blockIdx.x + G, ...
// and loads tile k + G into registers as soon as tile k is in LDS, so HBM
// reads continue while the block computes and stores.
template <bool SKEW>
__device__ __forceinline__ ulonglong2 tile_get1(const uint64_t *lds, int c) {
    const int e = c >> 5, k = (c & 31) * 2;
    ulonglong2 v;
    if (SKEW) {
        v.x = lds[e * SROW + skew(k)];
        v.y = lds[e * SROW + skew(k + 1)];
    } else {
        v = *reinterpret_cast<const ulonglong2 *>(&lds[e * SROW + k]);
    }
    return v;
}

// DPP: the exchange between the register phases through quad permutes
// (xch_a_to_b / xch_b_to_a) instead of LDS, which saves one barrier and 16
// LDS accesses per thread and tile for 64 VALU moves.
template <bool INV, bool DPP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_ntt_inc(
    uint64_t *__restrict__ data, size_t count, const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[EL * SROW];
    // twiddles in LDS: a twiddle read from global memory would have to wait
    // for the in-order vector-memory counter, i.e. for the prefetch as well
    __shared__ double twl[33];
    const size_t ntiles = (count + EL - 1) / EL;
    size_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    const int tid = threadIdx.x, e = tid >> 3, h = tid & 7, p = h >> 2, t = h & 3;
    // balanced twiddles for the lazy butterflies
    if (tid < 32) twl[tid] = balance(tab[(INV ? TAB_INV32 : TAB_FWD32) + tid], M);
    if (tid == 32) twl[32] = balance(tab[TAB_NINV32], M);
    uint64_t *row = lds + e * SROW;
    // the first tile into LDS, the next one's loads in flight
    int cur = (int)min((size_t)EL, count - tile * EL);
    const uint64_t *src = data + tile * EL * kD;
    ulonglong2 r0 = tile_chunk(src, tid, cur), r1 = tile_chunk(src, tid + NT, cur),
               r2 = tile_chunk(src, tid + 2 * NT, cur), r3 = tile_chunk(src, tid + 3 * NT, cur);
    // INV reads the IncompleteNTT layout (skewed); the forward reads plain rows
    tile_put1<INV>(lds, r0, tid);
    tile_put1<INV>(lds, r1, tid + NT);
    tile_put1<INV>(lds, r2, tid + 2 * NT);
    tile_put1<INV>(lds, r3, tid + 3 * NT);
    __syncthreads();
    // Every load and store below is unconditional, so the wait before the
    // next tile's LDS write only covers its own loads: past the last tile the
    // prefetch re-reads element 0, and a store of a chunk past nel rewrites
    // chunk 0 of the tile with its own value.
    size_t next = tile + gridDim.x;
    int nxt = next < ntiles ? (int)min((size_t)EL, count - next * EL) : 1;
    src = next < ntiles ? data + next * EL * kD : data;
    r0 = tile_chunk(src, tid, nxt);
    r1 = tile_chunk(src, tid + NT, nxt);
    r2 = tile_chunk(src, tid + 2 * NT, nxt);
    r3 = tile_chunk(src, tid + 3 * NT, nxt);
    for (;;) {
        const bool live = e < cur;
        double v[8];
        if (!INV) {
            if (live) {
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = u2d(row[8 * k + 2 * t + p]);  // coeff 2(t+4k)+p
                ntt32_fwd_a_lazy(v, twl, M);
                if (DPP) {
                    xch_a_to_b(v, t);
                    ntt32_fwd_b_lazy(v, twl, t, M);
                }
            }
            __syncthreads();
            if (live && !DPP) {
#pragma unroll
                for (int k = 0; k < 8; k++) row[36 * p + skew(t + 4 * k)] = dbits(v[k]);
            }
            if (!DPP) __syncthreads();
            if (live) {
                if (!DPP) {
#pragma unroll
                    for (int x = 0; x < 8; x++) v[x] = bitsd(row[36 * p + skew(8 * t + x)]);
                    ntt32_fwd_b_lazy(v, twl, t, M);
                }
                // output entry 32 p + 8 t + x, skewed: same slots this thread read
#pragma unroll
                for (int x = 0; x < 8; x++) row[36 * p + skew(8 * t + x)] = d2u(v[x]);
            }
        } else {
            if (live) {
#pragma unroll
                for (int x = 0; x < 8; x++) v[x] = u2d(row[36 * p + skew(8 * t + x)]);
                ntt32_inv_b_lazy(v, twl, t, M);
                if (DPP) {
                    xch_b_to_a(v, t);
                    ntt32_inv_a_lazy(v, twl, twl[32], M);
                } else {
#pragma unroll
                    for (int x = 0; x < 8; x++) row[36 * p + skew(8 * t + x)] = dbits(v[x]);
                }
            }
            if (!DPP) {
                __syncthreads();
                if (live) {
#pragma unroll
                    for (int k = 0; k < 8; k++) v[k] = bitsd(row[36 * p + skew(t + 4 * k)]);
                    ntt32_inv_a_lazy(v, twl, twl[32], M);
                }
            }
            __syncthreads();
            if (live) {
#pragma unroll
                for (int k = 0; k < 8; k++) row[8 * k + 2 * t + p] = d2u(v[k]);
            }
        }
        __syncthreads();
        // store the tile (forward output skewed, inverse output plain)
        {
            ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(data + tile * EL * kD);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                int c = tid + NT * i;
                c = c < cur * 32 ? c : 0;
                dst[c] = tile_get1<!INV>(lds, c);
            }
        }
        if (next >= ntiles) break;
        __syncthreads();  // the stores read LDS that the next tile overwrites
        tile_put1<INV>(lds, r0, tid);
        tile_put1<INV>(lds, r1, tid + NT);
        tile_put1<INV>(lds, r2, tid + 2 * NT);
        tile_put1<INV>(lds, r3, tid + 3 * NT);
        __syncthreads();
        tile = next;
        cur = nxt;
        next = tile + gridDim.x;
        nxt = next < ntiles ? (int)min((size_t)EL, count - next * EL) : 1;
        src = next < ntiles ? data + next * EL * kD : data;
        r0 = tile_chunk(src, tid, nxt);
        r1 = tile_chunk(src, tid + NT, nxt);
        r2 = tile_chunk(src, tid + 2 * NT, nxt);
        r3 = tile_chunk(src, tid + 3 * NT, nxt);
    }
}

// Forward IncompleteNTT of whole tiles fused with the full-NTT digit planes
// of every tile (persistent, pipelined as k_ntt_inc).  Once a tile's
// IncompleteNTT is in LDS it is stored, and its 64 full-NTT slots are formed
// from the halves into a second LDS buffer in the wpos layout: slots s and
// s + 1 (s even) evaluate at zeta_s and -zeta_s (brv6(s + 1) = brv6(s) + 32),
// so with zj = zj_s they are E[zj] +- zeta_s O[zj], and thread (e, h) forms
// slots 8 h .. 8 h + 7 of element e.  The next tile then moves into the first
// buffer while the planes are written (store_planes) where ptiles[tile] says.
// Tiles of TEL ring elements (8 threads each).  TEL = 64: 69 KiB of LDS, two
// blocks of 512 threads per CU; TEL = 128: 138 KiB, one block of 1024.
// Either way four waves per SIMD, so up to 128 registers.
template <int TEL>
__global__ __launch_bounds__(8 * TEL) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_ntt_fwd_planes(
    uint64_t *__restrict__ data, size_t ntiles, const double *__restrict__ tab, Mod M,
    int8_t *__restrict__ planes, const PlaneTile *__restrict__ ptiles) {
    constexpr int TNT = 8 * TEL;
    __shared__ uint64_t lds[TEL * SROW];
    __shared__ uint64_t work[64 * TEL];
    __shared__ double twl[32];
    __shared__ double zt[32];
    __shared__ int zo[32];
    size_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    const int tid = threadIdx.x, e = tid >> 3, h = tid & 7, p = h >> 2, t = h & 3;
    // balanced twiddles for the lazy butterflies
    if (tid < 32) twl[tid] = balance(tab[TAB_FWD32 + tid], M);
    if (tid >= 64 && tid < 96) {
        zt[tid - 64] = tab[TAB_ZETA + 2 * (tid - 64)];
        zo[tid - 64] = skew((int)tab[TAB_ZJ + 2 * (tid - 64)]);
    }
    uint64_t *row = lds + e * SROW;
    const uint64_t *src = data + tile * TEL * kD;
    ulonglong2 r0 = tile_chunk(src, tid, TEL), r1 = tile_chunk(src, tid + TNT, TEL),
               r2 = tile_chunk(src, tid + 2 * TNT, TEL), r3 = tile_chunk(src, tid + 3 * TNT, TEL);
    PlaneTile dsc = ptiles[tile];
    tile_put1<false>(lds, r0, tid);
    tile_put1<false>(lds, r1, tid + TNT);
    tile_put1<false>(lds, r2, tid + 2 * TNT);
    tile_put1<false>(lds, r3, tid + 3 * TNT);
    __syncthreads();
    // Loads and stores are unconditional (past the last tile the prefetch
    // re-reads tile 0), so each wait covers only the loads it needs.
    size_t next = tile + gridDim.x;
    src = next < ntiles ? data + next * TEL * kD : data;
    r0 = tile_chunk(src, tid, TEL);
    r1 = tile_chunk(src, tid + TNT, TEL);
    r2 = tile_chunk(src, tid + 2 * TNT, TEL);
    r3 = tile_chunk(src, tid + 3 * TNT, TEL);
    PlaneTile ndsc = ptiles[next < ntiles ? next : 0];
    for (;;) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = u2d(row[8 * k + 2 * t + p]);  // coeff 2(t+4k)+p
        ntt32_fwd_a_lazy(v, twl, M);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; k++) row[36 * p + skew(t + 4 * k)] = dbits(v[k]);
        __syncthreads();
#pragma unroll
        for (int x = 0; x < 8; x++) v[x] = bitsd(row[36 * p + skew(8 * t + x)]);
        ntt32_fwd_b_lazy(v, twl, t, M);
#pragma unroll
        for (int x = 0; x < 8; x++) row[36 * p + skew(8 * t + x)] = d2u(v[x]);
        __syncthreads();
        {
            ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(data + tile * TEL * kD);
#pragma unroll
            for (int i = 0; i < 4; i++) dst[tid + TNT * i] = tile_get1<true>(lds, tid + TNT * i);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int o = zo[4 * h + u];
            const double E = u2d(row[o]), O = u2d(row[36 + o]);
            const double z = mulmod(zt[4 * h + u], O, M);
            work[wpos<TEL>(8 * h + 2 * u, e)] = dbits(addmod(E, z, M));
            work[wpos<TEL>(8 * h + 2 * u + 1, e)] = dbits(submod(E, z, M));
        }
        __syncthreads();  // the tile rows are free and the slots complete
        const bool more = next < ntiles;
        if (more) {
            tile_put1<false>(lds, r0, tid);
            tile_put1<false>(lds, r1, tid + TNT);
            tile_put1<false>(lds, r2, tid + 2 * TNT);
            tile_put1<false>(lds, r3, tid + 3 * TNT);
        }
        store_planes_one<TEL>(planes + dsc.base, dsc.sstride, dsc.kstride, work, TEL, tid);
        if (!more) break;
        __syncthreads();  // the next tile is in LDS; the slot buffer has been read
        tile = next;
        dsc = ndsc;
        next = tile + gridDim.x;
        src = next < ntiles ? data + next * TEL * kD : data;
        r0 = tile_chunk(src, tid, TEL);
        r1 = tile_chunk(src, tid + TNT, TEL);
        r2 = tile_chunk(src, tid + 2 * TNT, TEL);
        r3 = tile_chunk(src, tid + 3 * TNT, TEL);
        ndsc = ptiles[next < ntiles ? next : 0];
    }
}

// The batched step's NTT pass with the transforms per wave: wave w of a
// block owns ring elements 8 w .. 8 w + 7 of the block's 128-element tile
// (4 KiB contiguous) with its own LDS rows and its own register prefetch of
// its next slice, so the transforms need only wave-level ordering and the 16
// waves of a CU overlap them.  Per tile: forward IncompleteNTT; the 64
// full-NTT slots of each element (slots s, s + 1 at +-zeta_s from E[zj],
// O[zj]) into the block's slot staging (wpos layout); two block barriers
// around it, after which every thread packs one slot x 8 columns into digit
// planes (store_planes: a wave's stores cover 4 slots, so each touches few
// pages); then, per wave again, the inverse IncompleteNTT from the forward
// outputs still in registers and the Coefficient slice stored.
static_assert(kPlaneTile == EL, "the round trip works on 128-element tiles");
constexpr int WEL = 8;                // ring elements per wave
constexpr int WROWS = WEL * SROW;     // u64 of a wave's rows
// DPP: both exchanges between register phases (forward a -> b, inverse
// b -> a) through quad permutes instead of the wave's LDS rows.
template <bool SHORT, bool DPP>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_ntt_rt_planes(
    uint64_t *__restrict__ data, size_t ntiles, const double *__restrict__ tab, Mod M,
    int8_t *__restrict__ planes, const PlaneTile *__restrict__ ptiles, uint16_t *__restrict__ sht,
    uint8_t *__restrict__ sht8, uint8_t *__restrict__ shfmt) {
    __shared__ uint64_t rows_all[(NT / 64) * WROWS];
    __shared__ uint64_t work[64 * EL];
    __shared__ double twl[32], twi[33], zt[32];
    __shared__ int zo[32];
    size_t tile = blockIdx.x;
    if (tile >= ntiles) return;  // uniform in the block
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ew = lane >> 3, h = lane & 7, p = h >> 2, t = h & 3;
    // balanced twiddles for the lazy butterflies; zeta_s and the skewed
    // position of E[zj_s] for even s
    if (tid < 32) twl[tid] = balance(tab[TAB_FWD32 + tid], M);
    if (tid >= 64 && tid < 96) twi[tid - 64] = balance(tab[TAB_INV32 + tid - 64], M);
    if (tid == 96) twi[32] = balance(tab[TAB_NINV32], M);
    if (tid >= 128 && tid < 160) {
        zt[tid - 128] = tab[TAB_ZETA + 2 * (tid - 128)];
        zo[tid - 128] = skew((int)tab[TAB_ZJ + 2 * (tid - 128)]);
    }
    __syncthreads();
    uint64_t *rows = rows_all + w * WROWS;
    uint64_t *row = rows + ew * SROW;
    const int e = WEL * w + ew;  // element of the block's tile
    // chunk c (16 B) of the slice: element c / 32, entries 2 (c % 32) ..
    auto slice = [&](size_t tl) { return data + (tl * EL + WEL * w) * kD; };
    const ulonglong2 *src = reinterpret_cast<const ulonglong2 *>(slice(tile));
    ulonglong2 r0 = src[lane], r1 = src[lane + 64], r2 = src[lane + 128], r3 = src[lane + 192];
    PlaneTile dsc = ptiles[tile];
    const uint64_t qi = (uint64_t)M.q, half = qi >> 1, off_neg = (uint64_t)kShortOff - qi;
    // the slice of tile tl into the wave's rows; with a short copy, also its
    // entries (chunk c holds coefficients 2c, 2c + 1): one byte each when the
    // whole slice is byte-short (one u16 store per chunk), else two bytes each
    // when it is short (one u32 store), and its format byte
    auto put = [&](size_t tl) {
        *reinterpret_cast<ulonglong2 *>(&rows[(lane >> 5) * SROW + (lane & 31) * 2]) = r0;
        *reinterpret_cast<ulonglong2 *>(&rows[(2 + (lane >> 5)) * SROW + (lane & 31) * 2]) = r1;
        *reinterpret_cast<ulonglong2 *>(&rows[(4 + (lane >> 5)) * SROW + (lane & 31) * 2]) = r2;
        *reinterpret_cast<ulonglong2 *>(&rows[(6 + (lane >> 5)) * SROW + (lane & 31) * 2]) = r3;
        if (SHORT) {
            const ulonglong2 rr[4] = {r0, r1, r2, r3};
            const size_t sl = tl * EL + WEL * w;  // the slice's first ring element
            uint32_t ts[4];
            uint32_t bad = 0, bad8 = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint64_t ta = short_t(rr[i].x, half, off_neg), tb = short_t(rr[i].y, half, off_neg);
                bad |= short_bad(ta) | short_bad(tb);
                bad8 |= byte_bad(ta) | byte_bad(tb);
                ts[i] = ((uint32_t)ta & 0xFFFFu) | ((uint32_t)tb << 16);
            }
            const bool any = __builtin_amdgcn_ballot_w64(bad != 0) != 0;
            const bool any8 = __builtin_amdgcn_ballot_w64(bad8 != 0) != 0;
            if (!any8) {
                // t - 2^14 + 128 = b + 128 in each byte
                uint16_t *dst = reinterpret_cast<uint16_t *>(sht8 + sl * kD) + lane;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint32_t z = ts[i] - ((kShortOff - 128u) | ((kShortOff - 128u) << 16));
                    dst[64 * i] = (uint16_t)((z & 0xFFu) | ((z >> 8) & 0xFF00u));
                }
            } else if (!any) {
                uint32_t *dst = reinterpret_cast<uint32_t *>(sht + sl * kD) + lane;
#pragma unroll
                for (int i = 0; i < 4; i++) dst[64 * i] = ts[i];
            }
            if (lane == 0) shfmt[sl * kD / kShortSlice] = any ? 2 : any8 ? 1 : 0;
        }
    };
    put(tile);
    wave_sync();
    // Loads and stores are unconditional (the last prefetch re-reads the
    // current slice), so each wait covers only the loads it needs.
    size_t next = tile + gridDim.x;
    src = reinterpret_cast<const ulonglong2 *>(slice(next < ntiles ? next : tile));
    r0 = src[lane];
    r1 = src[lane + 64];
    r2 = src[lane + 128];
    r3 = src[lane + 192];
    PlaneTile ndsc = ptiles[next < ntiles ? next : tile];
    for (;;) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = u2d(row[8 * k + 2 * t + p]);  // coeff 2(t+4k)+p
        ntt32_fwd_a_lazy(v, twl, M);
        if (DPP) {
            xch_a_to_b(v, t);
        } else {
            wave_sync();
#pragma unroll
            for (int k = 0; k < 8; k++) row[36 * p + skew(t + 4 * k)] = dbits(v[k]);
            wave_sync();
#pragma unroll
            for (int x = 0; x < 8; x++) v[x] = bitsd(row[36 * p + skew(8 * t + x)]);
        }
        ntt32_fwd_b_lazy(v, twl, t, M);
        if (DPP) wave_sync();  // every lane has read its coefficients
#pragma unroll
        for (int x = 0; x < 8; x++) row[36 * p + skew(8 * t + x)] = d2u(v[x]);
        wave_sync();
        // slots 8 h .. 8 h + 7 of element e
        double a[8];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int o = zo[4 * h + u];
            const double E = u2d(row[o]), O = u2d(row[36 + o]);
            const double z = mulmod(zt[4 * h + u], O, M);
            a[2 * u] = addmod(E, z, M);
            a[2 * u + 1] = submod(E, z, M);
        }
        ntt32_inv_b_lazy(v, twi, t, M);  // v: this lane's forward outputs 32p + 8t + x
        __syncthreads();  // the previous tile's planes have been read out of work
#pragma unroll
        for (int x = 0; x < 8; x++) work[wpos_w(8 * h + x, e)] = dbits(a[x]);
        __syncthreads();
        store_planes_one<EL, true>(planes + dsc.base, dsc.sstride, dsc.kstride, work, EL, tid);
        if (DPP) {
            xch_b_to_a(v, t);
        } else {
#pragma unroll
            for (int x = 0; x < 8; x++) row[36 * p + skew(8 * t + x)] = dbits(v[x]);
            wave_sync();
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bitsd(row[36 * p + skew(t + 4 * k)]);
        }
        ntt32_inv_a_lazy(v, twi, twi[32], M);
        wave_sync();
#pragma unroll
        for (int k = 0; k < 8; k++) row[8 * k + 2 * t + p] = d2u(v[k]);
        wave_sync();
        {
            ulonglong2 *dst = reinterpret_cast<ulonglong2 *>(slice(tile));
#pragma unroll
            for (int i = 0; i < 4; i++)
                dst[lane + 64 * i] =
                    *reinterpret_cast<const ulonglong2 *>(&rows[(2 * i + (lane >> 5)) * SROW + (lane & 31) * 2]);
        }
        if (next >= ntiles) break;
        wave_sync();  // the store has read the rows
        put(next);
        wave_sync();
        tile = next;
        dsc = ndsc;
        next = tile + gridDim.x;
        src = reinterpret_cast<const ulonglong2 *>(slice(next < ntiles ? next : tile));
        r0 = src[lane];
        r1 = src[lane + 64];
        r2 = src[lane + 128];
        r3 = src[lane + 192];
        ndsc = ptiles[next < ntiles ? next : tile];
    }
}

template <bool INV>
__global__ __launch_bounds__(NT) void k_ntt_full(uint64_t *__restrict__ data, size_t count,
                                                 const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[EL * SROW];
    const size_t e0 = (size_t)blockIdx.x * EL;
    const int nel = (int)min((size_t)EL, count - e0);
    const int tid = threadIdx.x, e = tid >> 3, t = tid & 7;
    const bool live = e < nel;
    uint64_t *row = lds + e * SROW;
    double v[8];
    if (!INV) {
        load_block<false>(lds, data + e0 * kD, nel, tid);
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = u2d(row[t + 8 * k]);
            ntt64_fwd_a(v, tab + TAB_FWD64, M);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) row[skew(t + 8 * k)] = dbits(v[k]);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int x = 0; x < 8; x++) v[x] = bitsd(row[skew(8 * t + x)]);
            ntt64_fwd_b(v, tab + TAB_FWD64, t, M);
#pragma unroll
            for (int x = 0; x < 8; x++) row[skew(8 * t + x)] = d2u(v[x]);
        }
        __syncthreads();
        store_block<true>(data + e0 * kD, lds, nel, tid);
    } else {
        load_block<true>(lds, data + e0 * kD, nel, tid);
        __syncthreads();
        if (live) {
#pragma unroll
            for (int x = 0; x < 8; x++) v[x] = u2d(row[skew(8 * t + x)]);
            ntt64_inv_b(v, tab + TAB_INV64, t, M);
#pragma unroll
            for (int x = 0; x < 8; x++) row[skew(8 * t + x)] = dbits(v[x]);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = bitsd(row[skew(t + 8 * k)]);
            ntt64_inv_a(v, tab + TAB_INV64, tab[TAB_NINV64], M);
        }
        __syncthreads();
        if (live) {
#pragma unroll
            for (int k = 0; k < 8; k++) row[t + 8 * k] = d2u(v[k]);
        }
        __syncthreads();
        store_block<false>(data + e0 * kD, lds, nel, tid);
    }
}

// Full forward NTT from phase-A inputs v (positions t + 8k); uses `xch` (a
// 128 x SROW region) for the exchange and leaves the 64 slots in
// out[wpos(s, e)] (64 x 128 u64; may alias xch).
__device__ __forceinline__ void full_fwd_to_work(double (&v)[8], uint64_t *xch, uint64_t *out,
                                                 int e, int t, bool live,
                                                 const double *__restrict__ tab, Mod M) {
    if (live) ntt64_fwd_a(v, tab + TAB_FWD64, M);
    __syncthreads();
    if (live) {
#pragma unroll
        for (int k = 0; k < 8; k++) xch[e * SROW + skew(t + 8 * k)] = dbits(v[k]);
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int x = 0; x < 8; x++) v[x] = bitsd(xch[e * SROW + skew(8 * t + x)]);
        ntt64_fwd_b(v, tab + TAB_FWD64, t, M);
    }
    __syncthreads();
    if (live) {
#pragma unroll
        for (int x = 0; x < 8; x++) out[wpos(8 * t + x, e)] = dbits(v[x]);
    }
    __syncthreads();
}

// slot-major store: dst_row(s)[i0 + e] for s < 64, e < nel
__device__ __forceinline__ void store_slotmajor(double *__restrict__ dst, const uint64_t *work,
                                                size_t rows, size_t row, size_t m, size_t i0,
                                                int nel, int tid) {
    for (int c = tid; c < 64 * EL; c += NT) {
        const int s = c >> 7, e = c & (EL - 1);
        if (e < nel) dst[((size_t)s * rows + row) * m + i0 + e] = bitsd(work[wpos(s, e)]);
    }
}

// grid (ceil(m / 128), rows).  Source row `row` starts at
// (row % row_lo) * stride_lo + (row / row_lo) * stride_hi (u64 units), which
// covers plain matrices and the halves of split (split.rs:21-28).
// PLANES = false: exact doubles dst[(s * rows + row) * m + i];
// PLANES = true: fragment-tiled int8 digit planes (store_planes) with
// `pstride` K-tiles of 64 columns.
// Full NTT of the (up to) 128 ring elements at base + i0 * d of one row into
// the row's slot-major doubles or digit planes.
template <bool PLANES>
__device__ __forceinline__ void full_sm_block(uint64_t *lds, const uint64_t *base, size_t m,
                                              size_t rows, size_t row, size_t i0, void *dst,
                                              size_t pstride, const double *__restrict__ tab,
                                              Mod M) {
    const int nel = i0 < m ? (int)min((size_t)EL, m - i0) : 0;
    const int tid = threadIdx.x, e = tid >> 3, t = tid & 7;
    load_block<false>(lds, base + i0 * kD, nel, tid);
    __syncthreads();
    double v[8];
    const bool live = e < nel;
    if (live) {
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = u2d(lds[e * SROW + t + 8 * k]);
    }
    full_fwd_to_work(v, lds, lds, e, t, live, tab, M);
    if (PLANES)
        store_planes(plane_base((int8_t *)dst, rows, row, i0), plane_sstride(rows, pstride),
                     plane_kstride(rows), lds, nel, tid);  // pstride = K-tiles
    else
        store_slotmajor((double *)dst, lds, rows, row, m, i0, nel, tid);
}

template <bool PLANES>
__global__ __launch_bounds__(NT) void k_ntt_full_sm(const uint64_t *__restrict__ src, size_t rows,
                                                    size_t m, size_t row_lo, size_t stride_lo,
                                                    size_t stride_hi, void *__restrict__ dst,
                                                    size_t pstride, const double *__restrict__ tab,
                                                    Mod M) {
    __shared__ uint64_t lds[EL * SROW];
    const size_t row = blockIdx.y;
    const uint64_t *base = src + (row % row_lo) * stride_lo + (row / row_lo) * stride_hi;
    full_sm_block<PLANES>(lds, base, m, rows, row, (size_t)blockIdx.x * EL, dst, pstride, tab, M);
}

// Batched: job j turns witness rows first .. first + rows - 1 (each its own
// length, zero beyond it) into one planes matrix of `ktiles` K tiles; block
// b of the job handles row b / (ktiles / 2), columns 128 (b % (ktiles / 2)).
__global__ __launch_bounds__(NT) void k_ntt_planes_batch(
    const uint64_t *__restrict__ arena, const unsigned long long *__restrict__ inst_off,
    const unsigned long long *__restrict__ inst_m, const PlaneJob *__restrict__ jobs, int njobs,
    const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[EL * SROW];
    const int b = blockIdx.x;
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const PlaneJob jb = jobs[lo];
    const int per_row = jb.ktiles / 2, local = b - jb.blk0;
    const int row = local / per_row, cb = local % per_row;
    const int inst = jb.first + row;
    full_sm_block<true>(lds, arena + inst_off[inst] * kD, inst_m[inst], jb.rows, row,
                        (size_t)cb * EL, jb.dst, jb.ktiles, tab, M);
}

// signed digit t of x (arithmetic.rs:1496-1570): x is split into pos (2x <= q)
// or neg (q - x); digit = (|.| >> t lg) & (radix - 1); limb = +-digit mod q
__device__ __forceinline__ uint64_t limb_of(uint64_t x, uint64_t q, unsigned lg, uint64_t mask,
                                            int tl) {
    const bool neg = 2 * x > q;
    const uint64_t v = neg ? q - x : x;
    const unsigned s = lg * (unsigned)tl;
    const uint64_t dg = (lg == 0) ? (v & mask) : (s >= 64 ? 0 : (v >> s) & mask);
    return neg ? (dg ? q - dg : 0) : dg;
}

// grid (ceil(m / 128), r).  Row rho of W (Coefficient) -> limb rows
// rho * parts + tl of new_W (Coefficient) and of dst (slot-major full NTT,
// rows = r * parts in total).
template <bool PLANES>
__global__ __launch_bounds__(NT) void k_decomp_sm(const uint64_t *__restrict__ W, size_t r,
                                                  size_t m, size_t parts,
                                                  const uint64_t *__restrict__ d_radix,
                                                  uint64_t *__restrict__ new_W,
                                                  void *__restrict__ dst, size_t pstride,
                                                  uint64_t q, const double *__restrict__ tab,
                                                  Mod M) {
    __shared__ uint64_t work[EL * SROW];
    const size_t rho = blockIdx.y;
    const size_t i0 = (size_t)blockIdx.x * EL;
    const int nel = i0 < m ? (int)min((size_t)EL, m - i0) : 0;
    const int tid = threadIdx.x, e = tid >> 3, t = tid & 7;
    const uint64_t radix = *d_radix;
    const unsigned lg = (unsigned)__ffsll((long long)radix) - 1u;
    const uint64_t mask = radix - 1;
    // coalesced load through LDS; each thread keeps the 8 coefficients it
    // transforms (positions t + 8k) in registers for all limbs, so one LDS
    // buffer serves the whole block and two blocks fit on a CU
    load_block<false>(work, W + (rho * m + i0) * kD, nel, tid);
    __syncthreads();
    const bool live = e < nel;
    uint64_t raw[8];
    if (live) {
#pragma unroll
        for (int k = 0; k < 8; k++) raw[k] = work[e * SROW + t + 8 * k];
    }
    __syncthreads();  // work becomes the NTT's exchange buffer
    const size_t rows = r * parts;
    for (int tl = 0; tl < (int)parts; tl++) {
        const size_t orow = rho * parts + tl;
        // limb witness row, Coefficient form: the 8 threads of an element
        // write its 64 coefficients as 8 runs of 64 contiguous bytes
        double v[8];
        if (live) {
            uint64_t *nw = new_W + (orow * m + i0 + e) * kD + t;
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t lv = limb_of(raw[k], q, lg, mask, tl);
                nw[8 * k] = lv;
                v[k] = u2d(lv);
            }
        }
        full_fwd_to_work(v, work, work, e, t, live, tab, M);
        if (PLANES)
            store_planes(plane_base((int8_t *)dst, rows, orow, i0),
                         plane_sstride(rows, pstride), plane_kstride(rows), work, nel, tid);
        else
            store_slotmajor((double *)dst, work, rows, orow, m, i0, nel, tid);
        __syncthreads();  // work is rewritten by the next limb
    }
}

// ---- the power-series CRS (compute_commitment_keys, crs.rs:43-72) on the device
//
// Row j of the CRS is m_j, m_j^2, ..., m_j^len.  In the full NTT domain each
// slot of a power is the slot's scalar power, so element i of row j is
// mhat_j[s]^(i + 1) slot by slot.  tpow[j][e][s] = mhat_j[s]^e for e < 128
// (k_crs_table); a block of the 128 elements i0 .. of row j takes
// b[s] = mhat_j[s]^(i0 + 1) once (square and multiply) and then one product
// per value, and writes the values where the key's layout wants them.

__device__ __forceinline__ double crs_pow(double b, unsigned long long e, Mod M) {
    double r = 1.0;
    while (e) {
        if (e & 1) r = mulmod(r, b, M);
        b = mulmod(b, b, M);
        e >>= 1;
    }
    return r;
}

__global__ void k_crs_table(const uint64_t *__restrict__ mhat, size_t rows,
                            uint64_t *__restrict__ tpow, Mod M) {
    const size_t total = rows * EL * kD;
    for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < total;
         x += (size_t)gridDim.x * blockDim.x) {
        const size_t j = x / (EL * kD);
        const int e = (int)((x / kD) % EL), s = (int)(x % kD);
        tpow[x] = d2u(crs_pow(u2d(mhat[j * kD + s]), (unsigned long long)e, M));
    }
}

// grid (ceil(m / 128), rows).  MODE 0: NTT-form rows dst[(j m + i) 64 + s];
// 1: the key's slot-major doubles; 2: the key's digit planes with `pstride`
// K tiles (as full_sm_block writes them from a Coefficient key).
template <int MODE>
__global__ __launch_bounds__(NT) void k_crs_sm(const uint64_t *__restrict__ mhat,
                                               const uint64_t *__restrict__ tpow, size_t rows,
                                               size_t m, void *__restrict__ dst, size_t pstride,
                                               Mod M) {
    __shared__ uint64_t lds[EL * SROW];
    __shared__ double bp[kD];
    const size_t row = blockIdx.y, i0 = (size_t)blockIdx.x * EL;
    const int nel = (int)min((size_t)EL, m - i0);
    const int tid = threadIdx.x, e = tid >> 3, t = tid & 7;
    if (tid < kD) bp[tid] = crs_pow(u2d(mhat[row * kD + tid]), (unsigned long long)(i0 + 1), M);
    __syncthreads();
    const uint64_t *tp = tpow + (row * EL + e) * kD;
    if (MODE == 0) {
        if (e < nel) {
            uint64_t *o = reinterpret_cast<uint64_t *>(dst) + (row * m + i0 + e) * kD;
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const int s = t + 8 * x;
                o[s] = d2u(mulmod(bp[s], u2d(tp[s]), M));
            }
        }
        return;
    }
    if (e < nel) {
#pragma unroll
        for (int x = 0; x < 8; x++) {
            const int s = 8 * t + x;
            lds[wpos(s, e)] = dbits(mulmod(bp[s], u2d(tp[s]), M));
        }
    }
    __syncthreads();
    if (MODE == 2)
        store_planes(plane_base((int8_t *)dst, rows, row, i0), plane_sstride(rows, pstride),
                     plane_kstride(rows), lds, nel, tid);
    else
        store_slotmajor((double *)dst, lds, rows, row, m, i0, nel, tid);
}

// ---- inner products with one power-series layer (parallel_dot_single_series_matrix)
//
// out[j] = sum_i layer[i] W[j][i] in R_q for up to DOT_MAXR rows: per tile of
// 128 elements the layer's 64-point NTT is taken once and kept in registers
// (thread (e, t): slots 8t .. 8t + 7 of element e), each row's tile is
// transformed and multiplied slot by slot, the products are summed over the
// wave's 8 elements with lane shuffles, over the waves in LDS and over the
// block's tiles in LDS, and each block adds its sums (canonical, so at most
// gridDim.x q < 2^63) to acc [r][64] with one atomic per slot.
constexpr int DOT_MAXR = 16;
__global__ __launch_bounds__(NT) void k_dot_series(const uint64_t *__restrict__ layer,
                                                   const uint64_t *__restrict__ W, size_t m, int r,
                                                   unsigned long long *__restrict__ acc,
                                                   const double *__restrict__ tab, Mod M) {
    __shared__ uint64_t lds[EL * SROW];
    __shared__ double wsum[NT / 64][64];
    __shared__ double bsum[DOT_MAXR][64];
    const int tid = threadIdx.x, e = tid >> 3, t = tid & 7, lane = tid & 63, wave = tid >> 6;
    for (int x = tid; x < r * 64; x += NT) bsum[x >> 6][x & 63] = 0.0;
    const size_t ntiles = (m + EL - 1) / EL;
    for (size_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const size_t i0 = tile * EL;
        const int nel = (int)min((size_t)EL, m - i0);
        const bool live = e < nel;
        auto ntt_tile = [&](const uint64_t *src, double(&v)[8]) {
            __syncthreads();  // the previous users of lds and wsum are done
            load_block<false>(lds, src + i0 * kD, nel, tid);
            __syncthreads();
            if (live) {
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = u2d(lds[e * SROW + t + 8 * k]);
            }
            full_fwd_to_work(v, lds, lds, e, t, live, tab, M);
            if (!live) {
#pragma unroll
                for (int x = 0; x < 8; x++) v[x] = 0.0;
            }
        };
        double lv[8];
        ntt_tile(layer, lv);
        for (int j = 0; j < r; j++) {
            double p[8];
            ntt_tile(W + (size_t)j * m * kD, p);
#pragma unroll
            for (int x = 0; x < 8; x++) {
                p[x] = mulmod(lv[x], p[x], M);
#pragma unroll
                for (int off = 8; off < 64; off <<= 1) p[x] = addmod(p[x], __shfl_xor(p[x], off, 64), M);
            }
            if (lane < 8) {
#pragma unroll
                for (int x = 0; x < 8; x++) wsum[wave][8 * lane + x] = p[x];
            }
            __syncthreads();
            if (tid < 64) {
                double sum = bsum[j][tid];
#pragma unroll
                for (int w = 0; w < NT / 64; w++) sum = addmod(sum, wsum[w][tid], M);
                bsum[j][tid] = sum;
            }
        }
    }
    __syncthreads();
    for (int x = tid; x < r * 64; x += NT)
        atomicAdd(&acc[x], (unsigned long long)d2u(bsum[x >> 6][x & 63]));
}

__global__ void k_dot_zero(unsigned long long *__restrict__ acc, int n) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) acc[x] = 0;
}

// slot sums acc[j][s] (mod q) -> out[j] in IncompleteNTT form (the commit's
// output map); one block of 64 threads per row
__global__ __launch_bounds__(64) void k_dot_finalize(const unsigned long long *__restrict__ acc,
                                                     uint64_t *__restrict__ out,
                                                     const double *__restrict__ tab, Mod M) {
    __shared__ double vs[64];
    const int j = blockIdx.x, tt = threadIdx.x;
    vs[tt] = u2d(acc[j * 64 + tt] % (unsigned long long)M.q);
    __syncthreads();
    const double *fin = tab + TAB_FIN;
    double a = 0.0;
#pragma unroll 8
    for (int s2 = 0; s2 < 64; s2++) {
        a += mul_lazy(fin[s2 * 64 + tt], vs[s2], M);
        if ((s2 & 3) == 3) a = reduce_lazy(a, M);
    }
    out[(size_t)j * kD + tt] = d2u(reduce(a, M));
}

}  // namespace

// resident blocks per CU of a kernel (occupancy API, cached)
template <bool INV, bool DPP>
int ntt_inc_occupancy() {
    static int occ = 0;
    if (occ == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void *>(&k_ntt_inc<INV, DPP>), NT, 0) != hipSuccess ||
            n < 1)
            n = 1;
        occ = n;
    }
    return occ;
}

template <bool INV, bool DPP>
void run_ntt_inc(Context *ctx, uint64_t *data, size_t count, unsigned grid, Mod M) {
    // persistent: every resident block slot once
    const size_t slots = (size_t)ctx->ncu * ntt_inc_occupancy<INV, DPP>();
    const unsigned g = (unsigned)std::min<size_t>(grid, slots);
    hipLaunchKernelGGL((k_ntt_inc<INV, DPP>), dim3(g), dim3(NT), 0, ctx->stream, data, count,
                       ctx->d_tables, M);
}

int launch_ntt(Context *ctx, uint64_t *data, size_t count, int mode, int inverse) {
    if (count == 0) return SALSAA_OK;
    Mod M{ctx->qd, ctx->qinv};
    const unsigned grid = ceil_div(count, EL);
    if (mode == SALSAA_REP_INCOMPLETE) {
        if (inverse)
            ctx->ntt_xch ? run_ntt_inc<true, true>(ctx, data, count, grid, M)
                         : run_ntt_inc<true, false>(ctx, data, count, grid, M);
        else
            ctx->ntt_xch ? run_ntt_inc<false, true>(ctx, data, count, grid, M)
                         : run_ntt_inc<false, false>(ctx, data, count, grid, M);
    } else {
        if (inverse)
            hipLaunchKernelGGL(k_ntt_full<true>, dim3(grid), dim3(NT), 0, ctx->stream, data, count,
                               ctx->d_tables, M);
        else
            hipLaunchKernelGGL(k_ntt_full<false>, dim3(grid), dim3(NT), 0, ctx->stream, data, count,
                               ctx->d_tables, M);
    }
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_ntt_fwd_planes(Context *ctx, uint64_t *data, size_t count, int8_t *planes,
                          const PlaneTile *tiles, int roundtrip, const ShortCopy *sc, int xch,
                          unsigned grid) {
    if (count == 0) return SALSAA_OK;
    SALSAA_CHECK(count % kPlaneTile == 0, "fused planes need whole tiles");
    SALSAA_CHECK(sc == nullptr || (roundtrip && sc->t && sc->t8 && sc->fmt),
                 "the short copy comes with the round trip");
    Mod M{ctx->qd, ctx->qinv};
    const size_t nt = count / kPlaneTile;
    static int occ[2] = {0, 0};
    static int occ_rt[2][2] = {{0, 0}, {0, 0}};  // [short][xch]
    const bool dpp = roundtrip && xch;
    const void *fn = !roundtrip ? reinterpret_cast<const void *>(&k_ntt_fwd_planes<kPlaneTile>)
                     : sc ? (dpp ? reinterpret_cast<const void *>(&k_ntt_rt_planes<true, true>)
                                 : reinterpret_cast<const void *>(&k_ntt_rt_planes<true, false>))
                          : (dpp ? reinterpret_cast<const void *>(&k_ntt_rt_planes<false, true>)
                                 : reinterpret_cast<const void *>(&k_ntt_rt_planes<false, false>));
    const int threads = roundtrip ? NT : 8 * kPlaneTile;
    int &o = roundtrip ? occ_rt[sc ? 1 : 0][dpp ? 1 : 0] : occ[0];
    if (o == 0) {
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, threads, 0) != hipSuccess || n < 1)
            n = 1;
        o = n;
    }
    const unsigned g = (unsigned)std::min<size_t>(nt, grid ? (size_t)grid : (size_t)ctx->ncu * o);
    if (roundtrip && sc && dpp)
        hipLaunchKernelGGL((k_ntt_rt_planes<true, true>), dim3(g), dim3(NT), 0, ctx->stream, data,
                           nt, ctx->d_tables, M, planes, tiles, sc->t, sc->t8, sc->fmt);
    else if (roundtrip && sc)
        hipLaunchKernelGGL((k_ntt_rt_planes<true, false>), dim3(g), dim3(NT), 0, ctx->stream, data,
                           nt, ctx->d_tables, M, planes, tiles, sc->t, sc->t8, sc->fmt);
    else if (roundtrip && dpp)
        hipLaunchKernelGGL((k_ntt_rt_planes<false, true>), dim3(g), dim3(NT), 0, ctx->stream, data,
                           nt, ctx->d_tables, M, planes, tiles, nullptr, nullptr, nullptr);
    else if (roundtrip)
        hipLaunchKernelGGL((k_ntt_rt_planes<false, false>), dim3(g), dim3(NT), 0, ctx->stream,
                           data, nt, ctx->d_tables, M, planes, tiles, nullptr, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_ntt_fwd_planes<kPlaneTile>), dim3(g), dim3(8 * kPlaneTile), 0,
                           ctx->stream, data, nt, ctx->d_tables, M, planes, tiles);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_ntt_full_to_slotmajor(Context *ctx, const uint64_t *src, size_t rows, size_t m,
                                 size_t row_lo, size_t stride_lo, size_t stride_hi,
                                 double *dst) {
    if (rows == 0 || m == 0) return SALSAA_OK;
    SALSAA_CHECK(rows < 65536, "too many rows for one launch");
    Mod M{ctx->qd, ctx->qinv};
    dim3 grid(ceil_div(m, EL), (unsigned)rows);
    hipLaunchKernelGGL(k_ntt_full_sm<false>, grid, dim3(NT), 0, ctx->stream, src, rows, m, row_lo,
                       stride_lo, stride_hi, (void *)dst, (size_t)0, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_ntt_full_to_planes(Context *ctx, const uint64_t *src, size_t rows, size_t m,
                              size_t row_lo, size_t stride_lo, size_t stride_hi, int8_t *dst,
                              size_t pstride) {
    if (rows == 0 || m == 0) return SALSAA_OK;
    SALSAA_CHECK(rows < 65536, "too many rows for one launch");
    SALSAA_CHECK(pstride == (m + EL - 1) / EL * 2, "K-tiles must be m rounded up to 128, over 64");
    Mod M{ctx->qd, ctx->qinv};
    dim3 grid(ceil_div(m, EL), (unsigned)rows);
    hipLaunchKernelGGL(k_ntt_full_sm<true>, grid, dim3(NT), 0, ctx->stream, src, rows, m, row_lo,
                       stride_lo, stride_hi, (void *)dst, pstride, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_ntt_planes_batch(Context *ctx, const uint64_t *arena, const unsigned long long *inst_off,
                            const unsigned long long *inst_m, const PlaneJob *jobs, size_t njobs,
                            size_t nblk) {
    SALSAA_CHECK(njobs >= 1 && nblk >= 1 && nblk < (1u << 31), "plane grid size out of range");
    Mod M{ctx->qd, ctx->qinv};
    hipLaunchKernelGGL(k_ntt_planes_batch, dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, arena,
                       inst_off, inst_m, jobs, (int)njobs, ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_decomp_to_slotmajor(Context *ctx, const uint64_t *W, size_t r, size_t m, size_t parts,
                               const uint64_t *d_radix, uint64_t *new_W, void *dst, int planes,
                               size_t pstride) {
    if (r == 0 || m == 0) return SALSAA_OK;
    SALSAA_CHECK(r < 65536, "too many rows for one launch");
    Mod M{ctx->qd, ctx->qinv};
    if (planes) {
        SALSAA_CHECK(pstride == (m + EL - 1) / EL * 2, "K-tiles must be m rounded up to 128, over 64");
        dim3 grid(ceil_div(m, EL), (unsigned)r);
        hipLaunchKernelGGL(k_decomp_sm<true>, grid, dim3(NT), 0, ctx->stream, W, r, m, parts,
                           d_radix, new_W, dst, pstride, ctx->q, ctx->d_tables, M);
    } else {
        dim3 grid(ceil_div(m, EL), (unsigned)r);
        hipLaunchKernelGGL(k_decomp_sm<false>, grid, dim3(NT), 0, ctx->stream, W, r, m, parts,
                           d_radix, new_W, dst, (size_t)0, ctx->q, ctx->d_tables, M);
    }
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_crs(Context *ctx, const uint64_t *module, size_t rows, size_t m, int mode, void *dst,
               size_t pstride) {
    SALSAA_CHECK(rows >= 1 && rows < 65536 && m >= 1, "CRS size out of range");
    SALSAA_CHECK(mode >= 0 && mode <= 2, "unknown CRS layout");
    SALSAA_CHECK(mode != 2 || pstride == (m + EL - 1) / EL * 2, "K-tiles must be m rounded up to 128, over 64");
    Mod M{ctx->qd, ctx->qinv};
    // scratch: the module in NTT form [rows][64], then tpow [rows][128][64]
    uint64_t *buf = nullptr;
    SALSAA_TRY(scratch_get(ctx, 11, (rows + rows * EL) * kD * sizeof(uint64_t), (void **)&buf));
    uint64_t *mhat = buf, *tpow = buf + rows * kD;
    SALSAA_HIP(hipMemcpyAsync(mhat, module, rows * kD * sizeof(uint64_t), hipMemcpyDeviceToDevice,
                              ctx->stream));
    SALSAA_TRY(launch_ntt(ctx, mhat, rows, SALSAA_REP_FULL, 0));
    const size_t nt = rows * EL * kD;
    hipLaunchKernelGGL(k_crs_table, dim3((unsigned)std::min<size_t>(ceil_div(nt, (size_t)256), 4096)),
                       dim3(256), 0, ctx->stream, mhat, rows, tpow, M);
    SALSAA_HIP(hipGetLastError());
    dim3 grid(ceil_div(m, EL), (unsigned)rows);
    if (mode == 0)
        hipLaunchKernelGGL(k_crs_sm<0>, grid, dim3(NT), 0, ctx->stream, mhat, tpow, rows, m, dst,
                           pstride, M);
    else if (mode == 1)
        hipLaunchKernelGGL(k_crs_sm<1>, grid, dim3(NT), 0, ctx->stream, mhat, tpow, rows, m, dst,
                           pstride, M);
    else
        hipLaunchKernelGGL(k_crs_sm<2>, grid, dim3(NT), 0, ctx->stream, mhat, tpow, rows, m, dst,
                           pstride, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

int launch_dot_series(Context *ctx, const uint64_t *layer, size_t m, const uint64_t *W, size_t r,
                      uint64_t *out) {
    SALSAA_CHECK(m >= 1 && r >= 1 && r <= (size_t)DOT_MAXR, "inner product size out of range");
    Mod M{ctx->qd, ctx->qinv};
    unsigned long long *acc = nullptr;
    SALSAA_TRY(scratch_get(ctx, 11, r * 64 * sizeof(unsigned long long), (void **)&acc));
    hipLaunchKernelGGL(k_dot_zero, dim3(ceil_div(r * 64, (size_t)256)), dim3(256), 0, ctx->stream,
                       acc, (int)(r * 64));
    const size_t ntiles = (m + EL - 1) / EL;
    const unsigned g = (unsigned)std::min<size_t>(ntiles, (size_t)ctx->ncu * 2);
    hipLaunchKernelGGL(k_dot_series, dim3(g), dim3(NT), 0, ctx->stream, layer, W, m, (int)r, acc,
                       ctx->d_tables, M);
    hipLaunchKernelGGL(k_dot_finalize, dim3((unsigned)r), dim3(64), 0, ctx->stream, acc, out,
                       ctx->d_tables, M);
    SALSAA_HIP(hipGetLastError());
    return SALSAA_OK;
}

}  // namespace salsaa
