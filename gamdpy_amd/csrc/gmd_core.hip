// gmd_core.hip -- kernels, launch drivers and the C ABI of gamdpy_amd.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "gamdpy_amd.h"
#include "gmd_kernels.h"

// ============================================================================
// Drivers
// ============================================================================
// Every step of a run inside one launch; the blocks of one system meet at a
// software barrier after each step, and the last to arrive runs the tail.
// a pointer that is the same in every lane, made visibly uniform (loads
// through it become scalar loads)
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (T *)(((unsigned long long)hi << 32) | lo);
}

template <int BASE, int CUT, int INTEG>
__device__ __forceinline__ void persistent_body(const DP *Pp, const BlockCtx &B) {
    extern __shared__ __align__(16) char lds[];
    PairP *pp = (PairP *)(lds + Pp->lds_pp);
    double *red = (double *)(lds + Pp->lds_red);
    char *build = lds + Pp->lds_build;
    load_pair_table(*Pp, pp);
    __syncthreads();
    const int nsteps = Pp->ctrl->nsteps;
    for (int s = 0; s <= nsteps; ++s) {
        const DP &P = *Pp;
        const bool rebuild = ld_flag(P.rebuild_flag + s) != 0;
        const StepInfo I = step_info(P, s, INTEG, rebuild);
        float acc[GMD_NRED];
#pragma unroll
        for (int c = 0; c < GMD_NRED; ++c) acc[c] = 0.f;
        step_body<BASE, CUT, INTEG>(P, B, I, pp, build, acc);
        if (I.need_ksum || I.need_mom || I.need_save) block_partials(P, B, acc, red);
        grid_barrier(P.ctrl, B.nb, [&]() { step_tail(P, I, red, B.nb); });
    }
}

// One step; the step index and the rebuild decision come from the control
// block that the previous step's tail wrote.
template <int BASE, int CUT, int INTEG>
__device__ __forceinline__ void step_once(const DP &P, const BlockCtx &B) {
    extern __shared__ __align__(16) char lds[];
    const int s = P.ctrl->step;
    if (s > P.ctrl->nsteps) return;
    const bool rebuild = P.ctrl->rebuild_now != 0;
    PairP *pp = (PairP *)(lds + P.lds_pp);
    double *red = (double *)(lds + P.lds_red);
    char *build = lds + P.lds_build;
    load_pair_table(P, pp);
    __syncthreads();
    const StepInfo I = step_info(P, s, INTEG, rebuild);
    float acc[GMD_NRED];
#pragma unroll
    for (int c = 0; c < GMD_NRED; ++c) acc[c] = 0.f;
    step_body<BASE, CUT, INTEG>(P, B, I, pp, build, acc);
    if (I.need_ksum || I.need_mom || I.need_save) block_partials(P, B, acc, red);
}

template <int BASE, int CUT, int INTEG>
__global__ void __launch_bounds__(256) k_persistent(const DP *__restrict__ Pp) {
    persistent_body<BASE, CUT, INTEG>(Pp, BlockCtx{(int)blockIdx.x, (int)gridDim.x});
}

// One step per launch.  The cross-block work (sums, thermostat, momentum
// shift, saver row, rebuild decision) is the separate one-block k_tail: a
// last-block-done scheme here would need an agent-scope release per block,
// each a write-back of the XCD's L2 under the blocks still running, which
// costs more than one extra launch.
template <int BASE, int CUT, int INTEG>
__global__ void __launch_bounds__(256, 4) k_step(const DP *__restrict__ Pp) {
    step_once<BASE, CUT, INTEG>(*Pp, BlockCtx{(int)blockIdx.x, (int)gridDim.x});
}

// Batched replicas: block b of the grid works for replica blk[b].rep as its
// block blk[b].bid of blk[b].nb.  Replicas never wait for each other: the
// persistent form synchronises only the blocks of one replica.
struct BatchBlk {
    int rep, bid, nb, pad;
};

template <int BASE, int CUT, int INTEG>
__global__ void __launch_bounds__(256) k_batch_persistent(const DP *const *Ps, const BatchBlk *blks) {
    const BatchBlk a = blks[blockIdx.x];
    const int rep = __builtin_amdgcn_readfirstlane(a.rep);
    if (rep < 0) return;
    const BlockCtx B = {__builtin_amdgcn_readfirstlane(a.bid), __builtin_amdgcn_readfirstlane(a.nb)};
    persistent_body<BASE, CUT, INTEG>(uniform_ptr(Ps[rep]), B);
}

template <int BASE, int CUT, int INTEG>
__global__ void __launch_bounds__(256) k_batch_step(const DP *const *Ps, const BatchBlk *blks) {
    const BatchBlk a = blks[blockIdx.x];
    const int rep = __builtin_amdgcn_readfirstlane(a.rep);
    if (rep < 0) return;
    const BlockCtx B = {__builtin_amdgcn_readfirstlane(a.bid), __builtin_amdgcn_readfirstlane(a.nb)};
    step_once<BASE, CUT, INTEG>(*uniform_ptr(Ps[rep]), B);
}

// The cross-block tail of a step as its own one-block kernel (the kernel
// boundary orders it after every block of k_step).
__global__ void __launch_bounds__(256) k_tail(const DP *__restrict__ Pp, int integrator) {
    const DP &P = *Pp;
    __shared__ double red[4 * GMD_NRED];
    const int s = P.ctrl->step;
    if (s > P.ctrl->nsteps) return;
    const StepInfo I = step_info(P, s, integrator, P.ctrl->rebuild_now != 0);
    step_tail(P, I, red, P.step_grid);
}

// the tails of all replicas of a batch: block r finishes replica r's step
__global__ void __launch_bounds__(256) k_batch_tail(const DP *const *Ps, int integrator) {
    const DP &P = *uniform_ptr(Ps[blockIdx.x]);
    __shared__ double red[4 * GMD_NRED];
    const int s = P.ctrl->step;
    if (s > P.ctrl->nsteps) return;
    const StepInfo I = step_info(P, s, integrator, P.ctrl->rebuild_now != 0);
    step_tail(P, I, red, P.step_grid);
}

// ---- run set-up: cell grid of the starting positions and the skin test ----
// Kernel parameters are generic pointers (an address-space qualifier would
// change the kernel's mangled name between the host and device passes).
__global__ void k_begin_run(Ctrl *Cp, int nsteps, float t0, int buf0, long long lstep0) {
    GMEM Ctrl *C = (GMEM Ctrl *)Cp;
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        C->nsteps = nsteps;
        C->t0 = t0;
        C->buf0 = buf0;
        C->lstep0 = lstep0;
        C->step = 0;
        C->red_count = 0;
        C->bar_count = 0;
        C->cell_overflow = 0;
    }
}

__global__ void __launch_bounds__(256) k_cells_clear(const DP *__restrict__ Pp) {
    const DP &P = *Pp;
    const int n = 3 * P.ncells;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x) P.ccount[c] = 0;
}

// skin test of the starting positions (nblist.py:58-86 at step 0) and, for
// linked lists, insertion into cell buffer 0
__global__ void __launch_bounds__(256) k_run_setup(const DP *__restrict__ Pp) {
    const DP &P = *Pp;
    cpos_t pos = P.ctrl->buf0 ? P.pos1 : P.pos0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.N; i += gridDim.x * blockDim.x) {
        const float4 r = pos[i], rr = P.rref[i];
        float dx, dy, dz;
        if (dist_sq_mic(r.x, r.y, r.z, rr.x, rr.y, rr.z, P.box0, P.box1, P.box2, dx, dy, dz) > P.skin2q)
            P.rebuild_flag[0] = 1;
        if (P.nblist_kind == 1) cell_insert(P, 0, i, r.x, r.y, r.z);
    }
}

__global__ void k_setup_finish(const DP *__restrict__ Pp) {
    const DP &P = *Pp;
    if (threadIdx.x == 0 && blockIdx.x == 0) P.ctrl->rebuild_now = P.rebuild_flag[0];
}

// ============================================================================
// Radial distribution (calculator_radial_distribution.py:76-109)
// ============================================================================
struct RdfP {
    int N, nt, bins, nclass, tile, ntiles;
    long long npairs;
    float L0, L1, L2, inv_bw;
    const GMEM float4 *box;    // per tile: (centre, -), (half extent, -); pairs of tiles farther apart than skip2 are skipped
    float skip2;
    cpos_t pos;
    const GMEM int *ptype;     // nullptr: type from pos.w
    const GMEM float2 *thr2;   // (thr[k], thr[k+1]) k = 0..bins-1; thr = bin thresholds on d^2
    float out_thr;             // smallest excluded d^2 (d >= min(L)/2)
    GMEM unsigned *part;       // [gridDim][nclass*bins]
};

// Unordered pairs i < j are visited once; each adds one count to the class
// {ti, tj}.  Ordered counts are recovered on the host (the MIC distance is
// symmetric bit for bit, so d(i,j) == d(j,i)).  The bin of a pair is found
// without a square root: a fast estimate k0 from v_sqrt_f32 is corrected
// against exact thresholds on d^2 (thr[k] = smallest float32 d^2 whose bin
// under the reference's float64 rule is >= k), so the counts are exact.
__global__ void __launch_bounds__(256) k_rdf(RdfP R) {
    extern __shared__ __align__(16) char lds[];
    const int nh = R.nclass * R.bins;
    unsigned *hist = (unsigned *)lds;
    float2 *thr = (float2 *)(lds + (((size_t)nh * 4 + 15) & ~(size_t)15));
    float4 *tj = (float4 *)(thr + R.bins);
    int *cls = (int *)(tj + R.tile);   // class offset table [nt*nt]
    for (int x = threadIdx.x; x < nh; x += blockDim.x) hist[x] = 0u;
    for (int x = threadIdx.x; x < R.bins; x += blockDim.x) thr[x] = R.thr2[x];
    for (int x = threadIdx.x; x < R.nt * R.nt; x += blockDim.x) {
        int a = x / R.nt, b = x % R.nt;
        int lo = min(a, b), hi = max(a, b);
        cls[x] = (lo * R.nt - (lo * (lo - 1)) / 2 + (hi - lo)) * R.bins;
    }
    __syncthreads();
    const float out_thr = R.out_thr;
    const long long T = R.ntiles;
    const int bmax = R.bins - 1;
    for (long long p = blockIdx.x; p < R.npairs; p += gridDim.x) {
        // decode p -> (I, J), J >= I, rows of length T - I
        double a = (double)(2 * T + 1);
        long long I = (long long)((a - sqrt(a * a - 8.0 * (double)p)) * 0.5);
        if (I < 0) I = 0;
        if (I > T - 1) I = T - 1;
        while (I > 0 && I * T - (I * (I - 1)) / 2 > p) --I;
        while (I + 1 < T && (I + 1) * T - ((I + 1) * I) / 2 <= p) ++I;
        const long long J = I + (p - (I * T - (I * (I - 1)) / 2));
        {
            // lower bound of the minimum-image distance between the two tiles
            const float4 ci = R.box[2 * I], hi = R.box[2 * I + 1], cj = R.box[2 * J], hj = R.box[2 * J + 1];
            const float gx = fmaxf(0.f, fabsf(mic1(ci.x - cj.x, R.L0)) - (hi.x + hj.x));
            const float gy = fmaxf(0.f, fabsf(mic1(ci.y - cj.y, R.L1)) - (hi.y + hj.y));
            const float gz = fmaxf(0.f, fabsf(mic1(ci.z - cj.z, R.L2)) - (hi.z + hj.z));
            if (gx * gx + gy * gy + gz * gz > R.skip2) continue;   // every pair is beyond L/2
        }
        __syncthreads();
        const int j0 = (int)(J * R.tile);
        for (int x = threadIdx.x; x < R.tile; x += blockDim.x) {
            const int j = j0 + x;
            if (j < R.N) {
                float4 rj = R.pos[j];
                if (R.ptype) rj.w = __int_as_float(R.ptype[j]);
                tj[x] = rj;
            }
        }
        __syncthreads();
        const int i = (int)(I * R.tile) + threadIdx.x;
        if (threadIdx.x < R.tile && i < R.N) {
            const float4 ri = R.pos[i];
            const int ti = R.ptype ? R.ptype[i] : __float_as_int(ri.w);
            const int *clsi = cls + ti * R.nt;
            const int jn = min(R.tile, R.N - j0);
            const int jstart = (I == J) ? (threadIdx.x + 1) : 0;
            // four pairs per iteration: the LDS reads and the dependent bin
            // arithmetic of the four overlap
            int jj = jstart;
            for (; jj + 4 <= jn; jj += 4) {
                float4 rj[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) rj[u] = tj[jj + u];
                float d2[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float dx, dy, dz;
                    d2[u] = dist_sq_mic(rj[u].x, rj[u].y, rj[u].z, ri.x, ri.y, ri.z, R.L0, R.L1, R.L2, dx, dy, dz);
                }
                int k[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) k[u] = min((int)(__builtin_amdgcn_sqrtf(d2[u]) * R.inv_bw), bmax);
                float2 th[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) th[u] = thr[k[u]];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int kk = k[u] - (d2[u] < th[u].x ? 1 : 0) + (d2[u] >= th[u].y ? 1 : 0);
                    if (d2[u] < out_thr) atomicAdd(&hist[clsi[__float_as_int(rj[u].w)] + kk], 1u);
                }
            }
            for (; jj < jn; ++jj) {
                const float4 rj = tj[jj];
                float dx, dy, dz;
                const float d2 = dist_sq_mic(rj.x, rj.y, rj.z, ri.x, ri.y, ri.z, R.L0, R.L1, R.L2, dx, dy, dz);
                if (d2 < out_thr) {
                    int k = min((int)(__builtin_amdgcn_sqrtf(d2) * R.inv_bw), bmax);
                    const float2 th = thr[k];
                    k = k - (d2 < th.x ? 1 : 0) + (d2 >= th.y ? 1 : 0);
                    atomicAdd(&hist[clsi[__float_as_int(rj.w)] + k], 1u);
                }
            }
        }
    }
    __syncthreads();
    GMEM unsigned *dst = R.part + (size_t)blockIdx.x * nh;
    for (int x = threadIdx.x; x < nh; x += blockDim.x) dst[x] = hist[x];
}

// Bounding box of each tile of 256 consecutive particles, in the frame of the
// tile's first particle (minimum-image displacements), so a tile that
// straddles the periodic boundary still gets a tight box.
__global__ void __launch_bounds__(256) k_rdf_boxes(const float4 *pos_p, int N, float L0, float L1, float L2,
                                                   float4 *box_p) {
    const GMEM float4 *pos = (const GMEM float4 *)pos_p;
    GMEM float4 *box = (GMEM float4 *)box_p;
    __shared__ float red[6][4];
    const int t = blockIdx.x, i0 = t * 256, i = i0 + threadIdx.x;
    const float4 r0 = pos[i0];
    float lo[3] = {0.f, 0.f, 0.f}, hi[3] = {0.f, 0.f, 0.f};
    if (i < N) {
        const float4 r = pos[i];
        const float d[3] = {mic1(r.x - r0.x, L0), mic1(r.y - r0.y, L1), mic1(r.z - r0.z, L2)};
        for (int k = 0; k < 3; ++k) lo[k] = hi[k] = d[k];
    }
    for (int k = 0; k < 3; ++k) {
        for (int o = 32; o > 0; o >>= 1) {
            lo[k] = fminf(lo[k], __shfl_xor(lo[k], o, 64));
            hi[k] = fmaxf(hi[k], __shfl_xor(hi[k], o, 64));
        }
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < 3; ++k) {
            red[k][w] = lo[k];
            red[3 + k][w] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        float c[3], h[3];
        for (int k = 0; k < 3; ++k) {
            float a = red[k][0], b = red[3 + k][0];
            for (int x = 1; x < 4; ++x) {
                a = fminf(a, red[k][x]);
                b = fmaxf(b, red[3 + k][x]);
            }
            h[k] = 0.5f * (b - a) * 1.0001f + 1e-4f;   // rounding margin
            c[k] = 0.5f * (a + b);
        }
        box[2 * t] = make_float4(r0.x + c[0], r0.y + c[1], r0.z + c[2], 0.f);
        box[2 * t + 1] = make_float4(h[0], h[1], h[2], 0.f);
    }
}

// Sum the per-block histograms: thread (x, chunk) adds 'per' blocks of bin x.
__global__ void k_rdf_reduce(const unsigned *part_p, int nblocks, int nh, int per, unsigned long long *out_p) {
    const GMEM unsigned *part = (const GMEM unsigned *)part_p;
    GMEM unsigned long long *out = (GMEM unsigned long long *)out_p;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nh) return;
    const int b0 = blockIdx.y * per, b1 = min(nblocks, b0 + per);
    unsigned long long s = 0;
    for (int b = b0; b < b1; ++b) s += part[(size_t)b * nh + x];
    if (s) atomicAdd(gen_ptr(out + x), s);
}

__global__ void k_pack_pos(const float *r, const int *ptype, int N, float4 *pos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) pos[i] = make_float4(r[3 * i], r[3 * i + 1], r[3 * i + 2], __int_as_float(ptype[i]));
}

// ============================================================================
// Renumbering of the particles in cell order (gather locality for the step,
// rebuild and histogram kernels).  Deterministic: cells in index order,
// particles of a cell in their previous order.
// ============================================================================
// Morton code of a cell's coordinates (bits interleaved x, y, z): cells close
// in space get close keys, so a run of consecutive particles is a compact
// block of cells
__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void __launch_bounds__(256) k_sort_count(const DP *__restrict__ Pp, int cur, int *key, int *cnt) {
    const DP &P = *Pp;
    cpos_t pos = cur ? P.pos1 : P.pos0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.N; i += gridDim.x * blockDim.x) {
        const float4 r = pos[i];
        const int c = cell_of(P, r.x, r.y, r.z);
        const unsigned cz = c % P.ncz, cy = (c / P.ncz) % P.ncy, cx = c / (P.ncz * P.ncy);
        const int k = (int)((spread3(cx) << 2) | (spread3(cy) << 1) | spread3(cz));
        key[i] = k;
        atomicAdd(cnt + k, 1);
    }
}

// exclusive prefix sum of a[0, n) in place (one block of 1024 threads)
__global__ void __launch_bounds__(1024) k_scan(int *a, int n) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (n + 1023) / 1024;
    const int b = min(n, t * per), e = min(n, b + per);
    int sum = 0;
    for (int k = b; k < e; ++k) sum += a[k];
    part[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = (t >= o) ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - sum;
    for (int k = b; k < e; ++k) {
        const int v = a[k];
        a[k] = run;
        run += v;
    }
}

__global__ void __launch_bounds__(256) k_sort_scatter(int N, const int *key, int *cursor, int *order) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        order[atomicAdd(cursor + key[i], 1)] = i;
}

// particles of one cell in ascending previous index
__global__ void __launch_bounds__(256) k_sort_cells(const int *start, int ncells, int N, int *order) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += gridDim.x * blockDim.x) {
        const int b = start[c], e = (c + 1 < ncells) ? start[c + 1] : N;
        for (int x = b + 1; x < e; ++x) {
            const int v = order[x];
            int y = x - 1;
            while (y >= b && order[y] > v) {
                order[y + 1] = order[y];
                --y;
            }
            order[y + 1] = v;
        }
    }
}

// new[k] = old[order[k]] for every per-particle array; the positions go to
// the other ping-pong buffer (the caller flips buf0)
struct PermBufs {
    float4 *vel, *frc, *rref, *obeta;
    float *sc;
    int *rim, *perm;
};

__global__ void __launch_bounds__(256) k_permute(const DP *__restrict__ Pp, int cur, const int *order, PermBufs T) {
    const DP &P = *Pp;
    const int N = P.N;
    cpos_t src = cur ? P.pos1 : P.pos0;
    pos_t dst = cur ? P.pos0 : P.pos1;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < N; k += gridDim.x * blockDim.x) {
        const int o = order[k];
        dst[k] = src[o];
        T.vel[k] = P.vel[o];
        T.frc[k] = P.frc[o];
        T.rref[k] = P.rref[o];
        T.obeta[k] = P.obeta[o];
        for (int c = 0; c < 5; ++c) T.sc[(size_t)c * N + k] = P.sc[(size_t)c * N + o];
        for (int c = 0; c < 3; ++c) T.rim[(size_t)c * N + k] = P.rim[(size_t)c * N + o];
        T.perm[k] = P.perm[o];
    }
}

// inverse of the permutation: order[perm[i]] = i (the caller's order back)
__global__ void __launch_bounds__(256) k_perm_inverse(const DP *__restrict__ Pp, int *order) {
    const DP &P = *Pp;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < P.N; i += gridDim.x * blockDim.x) order[P.perm[i]] = i;
}

// ============================================================================
// Host side
// ============================================================================
namespace {

thread_local std::string g_err;

int fail(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return -1;
}

#define HIPCHECK(x)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail("%s: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

template <typename T>
int dalloc(T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    HIPCHECK(hipMalloc((void **)p, n * sizeof(T)));
    return 0;
}

typedef void (*kfun)(const DP *);
typedef void (*bfun)(const DP *const *, const BatchBlk *);

// the kernel instances of one (pair function, cutoff, integrator) combination
struct Kernels {
    kfun pers, step;
    bfun bpers, bstep;
};

template <int BASE, int CUT, int INTEG>
Kernels pick() {
    return Kernels{k_persistent<BASE, CUT, INTEG>, k_step<BASE, CUT, INTEG>, k_batch_persistent<BASE, CUT, INTEG>,
                   k_batch_step<BASE, CUT, INTEG>};
}

template <int BASE, int CUT>
Kernels pick_integ(int integ) {
    switch (integ) {
        case 0: return pick<BASE, CUT, 0>();
        case 1: return pick<BASE, CUT, 1>();
        case 2: return pick<BASE, CUT, 2>();
        default: return pick<BASE, CUT, 3>();
    }
}

template <int BASE>
Kernels pick_cut(int cut, int integ) {
    switch (cut) {
        case 0: return pick_integ<BASE, 0>(integ);
        case 1: return pick_integ<BASE, 1>(integ);
        default: return pick_integ<BASE, 2>(integ);
    }
}

Kernels pick_kernels(int base, int cut, int integ) { return base == 0 ? pick_cut<0>(cut, integ) : pick_cut<1>(cut, integ); }

float i2f(int v) {
    float f;
    memcpy(&f, &v, 4);
    return f;
}

int f2i(float v) {
    int i;
    memcpy(&i, &v, 4);
    return i;
}

bool is_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }

int ilog2(int x) {
    int s = 0;
    while ((1 << s) < x) ++s;
    return s;
}

}  // namespace

static const int GRAPH_STEPS = 20;

struct gmd_ctx {
    gmd_desc d;
    std::vector<float> params;
    std::vector<int> excl;
    std::vector<float4> h_pos, h_vel;
    std::vector<int> h_rim;
    hipStream_t stream = nullptr;
    bool own_stream = true;   // false: a replica of a gmd_batch (the batch owns the stream and the context)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cus = 0;
    // device
    float4 *pos0 = nullptr, *pos1 = nullptr, *vel = nullptr, *frc = nullptr, *rref = nullptr, *obeta = nullptr;
    float *sc = nullptr;
    int *rim = nullptr, *nbl = nullptr, *nbc = nullptr;
    int *ccount = nullptr, *cmem = nullptr;   // cell grid, three rotating buffers
    int *d_excl = nullptr;
    float *d_params = nullptr;
    Ctrl *ctrl = nullptr;
    float *partial = nullptr;
    int *rebuild_flag = nullptr;
    int rebuild_cap = 0;
    float *saver_out = nullptr;
    size_t nbl_len = 0;
    // renumbering in cell order
    int *perm = nullptr, *skey = nullptr, *sorder = nullptr, *scnt = nullptr, *scur = nullptr;
    PermBufs tb = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int nkeys = 0;                // Morton key space of the renumbering
    int reorder_every = 0;        // steps between renumberings (0: never)
    long long since_reorder = 0;  // steps since the last renumbering
    bool permuted = false;        // perm is not the identity
    // state
    int buf0 = 0;
    long long lstep = 0;
    bool need_rebuild = true;   // forced rebuild pending (first build, gmd_force_rebuild)
    float last_ms = 0.f;
    // launch
    DP P;
    DP *d_P = nullptr;            // device copy of P (kernels take a pointer)
    Kernels kern = {nullptr, nullptr, nullptr, nullptr};
    int block = 256, grid_pers = 1, grid_step = 1, grid_aux = 1;
    size_t lds_pers = 0, lds_step = 0, lds_build = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
};

struct gmd_rdf {
    int N, nt, bins, nclass;
    float4 *pos = nullptr;
    float4 *tbox = nullptr;   // tile boxes
    int *ptype = nullptr;
    float *d_thr = nullptr;
    unsigned *part = nullptr;
    unsigned long long *d_out = nullptr;
    int grid = 0;
    float box[3] = {0, 0, 0};
    std::vector<float> thr;
    hipStream_t stream = nullptr;
};

extern "C" {

int gmd_abi_version(void) { return GMD_ABI_VERSION; }

int gmd_desc_size(void) { return (int)sizeof(gmd_desc); }

const char *gmd_last_error(void) { return g_err.c_str(); }

int gmd_get_device_info(gmd_device_info *out) {
    if (!out) return fail("gmd_get_device_info: null output");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHECK(hipGetDeviceProperties(&p, dev));
    out->compute_units = p.multiProcessorCount;
    out->wavefront_size = p.warpSize;
    out->max_threads_per_block = p.maxThreadsPerBlock;
    out->total_mem = (long long)p.totalGlobalMem;
    snprintf(out->name, sizeof out->name, "%s", p.name);
    snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    return 0;
}

static int ctx_free(gmd_ctx *c) {
    if (!c) return 0;
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    if (c->graph) hipGraphDestroy(c->graph);
    void *ptrs[] = {c->pos0, c->pos1, c->vel, c->frc, c->rref, c->obeta, c->sc, c->rim, c->nbl, c->nbc,
                    c->ccount, c->cmem, c->d_excl,
                    c->d_params, c->ctrl, c->partial, c->rebuild_flag, c->saver_out, c->d_P,
                    c->perm, c->skey, c->sorder, c->scnt, c->scur, c->tb.vel, c->tb.frc, c->tb.rref, c->tb.obeta,
                    c->tb.sc, c->tb.rim, c->tb.perm};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream && c->own_stream) hipStreamDestroy(c->stream);
    delete c;
    return 0;
}

static int ctx_create(const gmd_desc *desc, gmd_ctx **out, hipStream_t shared_stream);
static int unpermute(gmd_ctx *c);
static int host_perm(gmd_ctx *c, std::vector<int> &perm);

int gmd_create(const gmd_desc *desc, gmd_ctx **out) { return ctx_create(desc, out, nullptr); }

// shared_stream != nullptr: the context works on the caller's stream (a batch
// of replicas) instead of creating its own
static int ctx_create(const gmd_desc *desc, gmd_ctx **out, hipStream_t shared_stream) {
    if (!desc || !out) return fail("gmd_create: null argument");
    *out = nullptr;
    const gmd_desc &d = *desc;
    if (d.abi_version != GMD_ABI_VERSION)
        return fail("gmd_create: ABI version %d, library expects %d", d.abi_version, GMD_ABI_VERSION);
    if (d.D != 3) return fail("gmd_create: only D=3 is supported (got D=%d)", d.D);
    if (d.N < 1) return fail("gmd_create: N must be positive (got %d)", d.N);
    if (d.ntypes < 1 || d.ntypes > MAX_TYPES) return fail("gmd_create: ntypes must be in [1,%d]", MAX_TYPES);
    if (d.pair_base < 0 || d.pair_base > 1) return fail("gmd_create: unknown pair function %d", d.pair_base);
    if (d.cutoff < 0 || d.cutoff > 2) return fail("gmd_create: unknown cutoff wrapper %d", d.cutoff);
    if (d.nblist_kind < 0 || d.nblist_kind > 1) return fail("gmd_create: unknown nblist kind %d", d.nblist_kind);
    if (d.integrator < 0 || d.integrator > 3) return fail("gmd_create: unknown integrator %d", d.integrator);
    if (d.max_nbs < 1) return fail("gmd_create: max_nbs must be positive");
    if (!d.params) return fail("gmd_create: params missing");
    if (!is_pow2(d.tp) || d.tp > 64) return fail("gmd_create: tp must be a power of two <= 64 (got %d)", d.tp);
    if (d.pb < 1) return fail("gmd_create: pb must be positive");
    if (d.utilize_niii) return fail("gmd_create: UtilizeNIII=True is not supported by the MI355X path");
    for (int k = 0; k < 3; ++k)
        if (!(d.box[k] > 0.f)) return fail("gmd_create: box lengths must be positive");
    if (d.scalar_save_every > 0 && d.scalar_saves_per_block < 1)
        return fail("gmd_create: scalar_saves_per_block must be positive");

    gmd_ctx *c = new gmd_ctx();
    c->d = d;
    const int N = d.N;
    c->params.assign(d.params, d.params + (size_t)d.ntypes * d.ntypes * 3);
    if (d.exclusions && d.exclusions_cols > 0) c->excl.assign(d.exclusions, d.exclusions + (size_t)N * d.exclusions_cols);
    c->d.params = nullptr;
    c->d.exclusions = nullptr;

    int rc = 0;
#define TRY(x)                 \
    do {                       \
        if ((rc = (x)) != 0) { \
            ctx_free(c);       \
            return rc;         \
        }                      \
    } while (0)
#define HTRY(x)                                                                 \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            ctx_free(c);                                                        \
            return fail("%s: %s", #x, hipGetErrorString(e_));                   \
        }                                                                       \
    } while (0)

    int dev = 0;
    HTRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HTRY(hipGetDeviceProperties(&prop, dev));
    c->cus = prop.multiProcessorCount;
    if (shared_stream) {
        c->stream = shared_stream;
        c->own_stream = false;
    } else {
        HTRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    }
    HTRY(hipEventCreate(&c->ev0));
    HTRY(hipEventCreate(&c->ev1));

    // ---- cell grid (nblist_linked_lists.py:36-42) ----
    int ncx = 1, ncy = 1, ncz = 1;
    if (d.nblist_kind == GMD_NBLIST_LINKED_LISTS) {
        double min_size = ((double)d.max_cut + (double)d.skin) / 2.0;
        ncx = (int)((double)d.box[0] / min_size);
        ncy = (int)((double)d.box[1] / min_size);
        ncz = (int)((double)d.box[2] / min_size);
        if (ncx <= 6 || ncy <= 4 || ncz <= 4) {
            ctx_free(c);
            return fail("linked lists need more than 6 cells in x and more than 4 in y and z "
                        "(box/((max_cut+skin)/2) gives %d x %d x %d); use 'N squared'", ncx, ncy, ncz);
        }
    }
    const int ncells = ncx * ncy * ncz;
    if (d.nblist_kind == GMD_NBLIST_LINKED_LISTS && N >= LL_MAX_N) {
        ctx_free(c);
        return fail("linked lists support at most %d particles (got %d)", LL_MAX_N - 1, N);
    }

    // ---- launch geometry ----
    const int tp = d.tp;
    int block = d.pb * tp;
    if (block % 64) block = ((block + 63) / 64) * 64;
    if (block > 256) block = 256;
    c->block = block;
    const int slots = block / tp;
    c->grid_step = (N + slots - 1) / slots;
    c->grid_aux = std::max(1, std::min((N + block - 1) / block, c->cus * 16));
    const int wpb = block / 64;
    // dynamic LDS: pair table | reduction scratch (one row of sums per wave) |
    // rebuild scratch; sized so that four blocks of 256 fit in a CU's 160 KB
    const size_t lds_cells = (d.nblist_kind == GMD_NBLIST_LINKED_LISTS) ? (size_t)wpb * LL_WAVE_LDS : 0;
    c->lds_build = std::max(lds_cells, (size_t)block * 16);
    const size_t lds_pp = 0;
    const size_t lds_red = ((size_t)d.ntypes * d.ntypes * sizeof(PairP) + 15) / 16 * 16;
    const size_t lds_b = lds_red + (size_t)wpb * GMD_NRED * sizeof(double);
    c->lds_step = lds_b + c->lds_build;
    c->lds_pers = c->lds_step;
    // cell capacity: room for 4x the mean occupancy, at least 16, at most 64
    int ccap = 16;
    if (d.nblist_kind == GMD_NBLIST_LINKED_LISTS) {
        const double mean = (double)N / ncells;
        while (ccap < 4.0 * mean + 8.0 && ccap < LL_OWNCAP) ccap *= 2;
        if (4.0 * mean + 8.0 > LL_OWNCAP) {
            ctx_free(c);
            return fail("linked lists: %.1f particles per cell on average exceeds the cell capacity %d", mean, LL_OWNCAP);
        }
    }

    // ---- buffers ----
    // at least 8 rows: the force loop requests the first 8 rows before the count arrives
    const size_t rows = std::max((size_t)(d.max_nbs + tp - 1) / tp, (size_t)8);
    c->nbl_len = rows * (size_t)N * tp;
    TRY(dalloc(&c->pos0, N));
    TRY(dalloc(&c->pos1, N));
    TRY(dalloc(&c->vel, N));
    TRY(dalloc(&c->frc, N));
    TRY(dalloc(&c->rref, N));
    TRY(dalloc(&c->obeta, N));
    TRY(dalloc(&c->sc, (size_t)5 * N));
    TRY(dalloc(&c->rim, (size_t)3 * N));
    TRY(dalloc(&c->nbl, c->nbl_len));
    TRY(dalloc(&c->nbc, N));
    TRY(dalloc(&c->ccount, (size_t)3 * ncells));
    TRY(dalloc(&c->cmem, (size_t)3 * ncells * ccap));
    TRY(dalloc(&c->d_params, c->params.size()));
    TRY(dalloc(&c->ctrl, 1));
    const int max_grid = std::max(c->grid_step, c->cus * 32);
    TRY(dalloc(&c->partial, (size_t)max_grid * GMD_NRED));
    TRY(dalloc(&c->saver_out, (size_t)std::max(1, d.scalar_saves_per_block) * 6));
    TRY(dalloc(&c->perm, N));
    {
        std::vector<int> iota(N);
        for (int i = 0; i < N; ++i) iota[i] = i;
        HTRY(hipMemcpy(c->perm, iota.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice));
    }
    int side = 1;
    while (side < std::max(ncx, std::max(ncy, ncz))) side *= 2;
    if (d.nblist_kind == GMD_NBLIST_LINKED_LISTS && c->excl.empty() && side <= 256) {
        // renumbering buffers (the exclusion table is indexed by particle, so
        // systems with exclusions keep the caller's numbering)
        TRY(dalloc(&c->skey, N));
        TRY(dalloc(&c->sorder, N));
        c->nkeys = side * side * side;
        TRY(dalloc(&c->scnt, (size_t)c->nkeys));
        TRY(dalloc(&c->scur, (size_t)c->nkeys));
        TRY(dalloc(&c->tb.vel, N));
        TRY(dalloc(&c->tb.frc, N));
        TRY(dalloc(&c->tb.rref, N));
        TRY(dalloc(&c->tb.obeta, N));
        TRY(dalloc(&c->tb.sc, (size_t)5 * N));
        TRY(dalloc(&c->tb.rim, (size_t)3 * N));
        TRY(dalloc(&c->tb.perm, N));
        c->reorder_every = 400;
        c->since_reorder = 1LL << 40;
    }
    if (!c->excl.empty()) {
        TRY(dalloc(&c->d_excl, c->excl.size()));
        HTRY(hipMemcpy(c->d_excl, c->excl.data(), c->excl.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    HTRY(hipMemcpy(c->d_params, c->params.data(), c->params.size() * sizeof(float), hipMemcpyHostToDevice));
    HTRY(hipMemset(c->ccount, 0, (size_t)3 * ncells * sizeof(int)));
    HTRY(hipMemset(c->sc, 0, (size_t)5 * N * sizeof(float)));
    HTRY(hipMemset(c->rim, 0, (size_t)3 * N * sizeof(int)));
    HTRY(hipMemset(c->nbc, 0, (size_t)N * sizeof(int)));
    HTRY(hipMemset(c->nbl, 0, c->nbl_len * sizeof(int)));   // entries past a count stay valid indices
    HTRY(hipMemset(c->cmem, 0, (size_t)3 * ncells * ccap * sizeof(int)));
    HTRY(hipMemset(c->obeta, 0, (size_t)N * sizeof(float4)));
    HTRY(hipMemset(c->frc, 0, (size_t)N * sizeof(float4)));
    HTRY(hipMemset(c->rref, 0, (size_t)N * sizeof(float4)));
    HTRY(hipMemset(c->saver_out, 0, (size_t)std::max(1, d.scalar_saves_per_block) * 6 * sizeof(float)));
    Ctrl h_ctrl;
    memset(&h_ctrl, 0, sizeof h_ctrl);
    h_ctrl.shift_step = -1;
    HTRY(hipMemcpy(c->ctrl, &h_ctrl, sizeof h_ctrl, hipMemcpyHostToDevice));
    c->h_pos.assign(N, make_float4(0.f, 0.f, 0.f, 0.f));
    c->h_vel.assign(N, make_float4(0.f, 0.f, 0.f, 1.f));
    c->h_rim.assign((size_t)3 * N, 0);

    // ---- device parameter block ----
    DP &P = c->P;
    memset(&P, 0, sizeof P);
    P.N = N;
    P.ntypes = d.ntypes;
    P.tp = tp;
    P.tp_shift = ilog2(tp);
    P.max_nbs = d.max_nbs;
    P.nblist_kind = d.nblist_kind;
    P.niii = 0;
    P.box0 = d.box[0];
    P.box1 = d.box[1];
    P.box2 = d.box[2];
    {
        float cps = d.max_cut + d.skin;
        P.cps2 = cps * cps;
        P.skin2q = (d.skin * d.skin) * 0.25f;
    }
    P.pos0 = (GMEM float4 *)c->pos0;
    P.pos1 = (GMEM float4 *)c->pos1;
    P.vel = (GMEM float4 *)c->vel;
    P.frc = (GMEM float4 *)c->frc;
    P.rref = (GMEM float4 *)c->rref;
    P.obeta = (GMEM float4 *)c->obeta;
    P.sc = (GMEM float *)c->sc;
    P.rim = (GMEM int *)c->rim;
    P.nbl = (GMEM int *)c->nbl;
    P.nbc = (GMEM int *)c->nbc;
    P.ncx = ncx;
    P.ncy = ncy;
    P.ncz = ncz;
    P.ncells = ncells;
    P.ccap = ccap;
    P.ccount = (GMEM int *)c->ccount;
    P.cmem = (GMEM int *)c->cmem;
    P.excl = (const GMEM int *)c->d_excl;
    P.excl_cols = d.exclusions_cols;
    P.raw_params = (const GMEM float *)c->d_params;
    P.pair_base = d.pair_base;
    P.cutoff = d.cutoff;
    P.dt = d.dt;
    P.omega2 = d.omega2;
    P.alpha = d.alpha;
    P.degrees = (double)d.degrees;
    P.seed = d.seed;
    P.tfunc = d.tfunc;
    P.tpar0 = d.tparams[0];
    P.tpar1 = d.tparams[1];
    P.tpar2 = d.tparams[2];
    P.tpar3 = d.tparams[3];
    P.compute_u = d.compute_u;
    P.compute_w = d.compute_w;
    P.compute_lap = d.compute_lap;
    P.compute_k = d.compute_k;
    P.compute_fsq = d.compute_fsq;
    P.mom_every = d.momentum_reset_every;
    P.save_every = d.scalar_save_every;
    P.saves_per_block = d.scalar_saves_per_block;
    P.ctrl = (GMEM Ctrl *)c->ctrl;
    P.partial = (GMEM float *)c->partial;
    P.pstride = max_grid;
    P.saver_out = (GMEM float *)c->saver_out;
    P.lds_pp = (int)lds_pp;
    P.lds_red = (int)lds_red;
    P.lds_build = (int)lds_b;
    P.step_grid = c->grid_step;
    P.perm = (const GMEM int *)c->perm;

    TRY(dalloc(&c->d_P, 1));
    c->kern = pick_kernels(d.pair_base, d.cutoff, d.integrator);

    // ---- cooperative grid: all blocks must be co-resident ----
    if (d.gridsync) {
        int nb = 0;
        HTRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)c->kern.pers, block, c->lds_pers));
        if (nb < 1) {
            ctx_free(c);
            return fail("gridsync: persistent kernel cannot be resident with %d threads per block", block);
        }
        c->grid_pers = std::max(1, std::min(c->grid_step, nb * c->cus));
    }
    *out = c;
    return 0;
#undef TRY
#undef HTRY
}

int gmd_destroy(gmd_ctx *ctx) {
    if (ctx && !ctx->own_stream) return fail("gmd_destroy: this context belongs to a batch; use gmd_batch_destroy");
    return ctx_free(ctx);
}

int gmd_upload(gmd_ctx *c, const float *r, const float *v, const float *f, const float *scalars, const int *ptype,
               const int *r_im) {
    if (!c) return fail("gmd_upload: null context");
    if (unpermute(c)) return -1;   // host arrays come in the caller's numbering
    const int N = c->d.N;
    const gmd_desc &d = c->d;
    if (ptype)
        for (int i = 0; i < N; ++i) {
            if (ptype[i] < 0 || ptype[i] >= d.ntypes)
                return fail("gmd_upload: particle %d has type %d outside [0,%d)", i, ptype[i], d.ntypes);
            c->h_pos[i].w = i2f(ptype[i]);
        }
    if (r)
        for (int i = 0; i < N; ++i) {
            c->h_pos[i].x = r[3 * i];
            c->h_pos[i].y = r[3 * i + 1];
            c->h_pos[i].z = r[3 * i + 2];
        }
    if (v)
        for (int i = 0; i < N; ++i) {
            c->h_vel[i].x = v[3 * i];
            c->h_vel[i].y = v[3 * i + 1];
            c->h_vel[i].z = v[3 * i + 2];
        }
    if (scalars && d.sid_m >= 0)
        for (int i = 0; i < N; ++i) c->h_vel[i].w = scalars[(size_t)i * d.nscalars + d.sid_m];
    if (r || ptype) {
        // the current buffer becomes pos0; whether the list is still valid is
        // decided by the skin test at the start of the next run
        c->buf0 = 0;
        HIPCHECK(hipMemcpy(c->pos0, c->h_pos.data(), (size_t)N * sizeof(float4), hipMemcpyHostToDevice));
    }
    if (v || (scalars && d.sid_m >= 0))
        HIPCHECK(hipMemcpy(c->vel, c->h_vel.data(), (size_t)N * sizeof(float4), hipMemcpyHostToDevice));
    if (f) {
        std::vector<float4> h(N);
        for (int i = 0; i < N; ++i) h[i] = make_float4(f[3 * i], f[3 * i + 1], f[3 * i + 2], 0.f);
        HIPCHECK(hipMemcpy(c->frc, h.data(), (size_t)N * sizeof(float4), hipMemcpyHostToDevice));
    }
    if (scalars) {
        std::vector<float> h((size_t)5 * N, 0.f);
        const int sids[5] = {d.sid_u, d.sid_w, d.sid_lap, d.sid_k, d.sid_fsq};
        for (int col = 0; col < 5; ++col)
            if (sids[col] >= 0)
                for (int i = 0; i < N; ++i) h[(size_t)col * N + i] = scalars[(size_t)i * d.nscalars + sids[col]];
        HIPCHECK(hipMemcpy(c->sc, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    if (r_im) {
        for (int i = 0; i < N; ++i)
            for (int k = 0; k < 3; ++k) c->h_rim[(size_t)k * N + i] = r_im[3 * i + k];
        HIPCHECK(hipMemcpy(c->rim, c->h_rim.data(), c->h_rim.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    return 0;
}

int gmd_download(gmd_ctx *c, float *r, float *v, float *f, float *scalars, int *ptype, int *r_im) {
    if (!c) return fail("gmd_download: null context");
    const int N = c->d.N;
    const gmd_desc &d = c->d;
    HIPCHECK(hipStreamSynchronize(c->stream));
    std::vector<int> perm;   // internal index -> caller's index (empty: identity)
    if (host_perm(c, perm)) return -1;
    auto at = [&](int i) { return perm.empty() ? i : perm[i]; };
    if (r || ptype) {
        HIPCHECK(hipMemcpy(c->h_pos.data(), c->buf0 ? c->pos1 : c->pos0, (size_t)N * sizeof(float4),
                           hipMemcpyDeviceToHost));
        for (int i = 0; i < N; ++i) {
            const int o = at(i);
            if (r) {
                r[3 * o] = c->h_pos[i].x;
                r[3 * o + 1] = c->h_pos[i].y;
                r[3 * o + 2] = c->h_pos[i].z;
            }
            if (ptype) ptype[o] = f2i(c->h_pos[i].w);
        }
    }
    if (v || scalars) {
        HIPCHECK(hipMemcpy(c->h_vel.data(), c->vel, (size_t)N * sizeof(float4), hipMemcpyDeviceToHost));
        if (v)
            for (int i = 0; i < N; ++i) {
                const int o = at(i);
                v[3 * o] = c->h_vel[i].x;
                v[3 * o + 1] = c->h_vel[i].y;
                v[3 * o + 2] = c->h_vel[i].z;
            }
    }
    if (f) {
        std::vector<float4> h(N);
        HIPCHECK(hipMemcpy(h.data(), c->frc, (size_t)N * sizeof(float4), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; ++i) {
            const int o = at(i);
            f[3 * o] = h[i].x;
            f[3 * o + 1] = h[i].y;
            f[3 * o + 2] = h[i].z;
        }
    }
    if (scalars) {
        std::vector<float> h((size_t)5 * N);
        HIPCHECK(hipMemcpy(h.data(), c->sc, h.size() * sizeof(float), hipMemcpyDeviceToHost));
        const int sids[5] = {d.sid_u, d.sid_w, d.sid_lap, d.sid_k, d.sid_fsq};
        for (int col = 0; col < 5; ++col)
            if (sids[col] >= 0)
                for (int i = 0; i < N; ++i) scalars[(size_t)at(i) * d.nscalars + sids[col]] = h[(size_t)col * N + i];
        if (d.sid_m >= 0)
            for (int i = 0; i < N; ++i) scalars[(size_t)at(i) * d.nscalars + d.sid_m] = c->h_vel[i].w;
    }
    if (r_im) {
        HIPCHECK(hipMemcpy(c->h_rim.data(), c->rim, c->h_rim.size() * sizeof(int), hipMemcpyDeviceToHost));
        for (int i = 0; i < N; ++i)
            for (int k = 0; k < 3; ++k) r_im[3 * at(i) + k] = c->h_rim[(size_t)k * N + i];
    }
    return 0;
}

static int ensure_flags(gmd_ctx *c, int steps) {
    if (c->rebuild_cap < steps + 2) {
        if (c->rebuild_flag) HIPCHECK(hipFree(c->rebuild_flag));
        c->rebuild_flag = nullptr;
        int cap = std::max(steps + 2, 1024);
        if (dalloc(&c->rebuild_flag, cap)) return -1;
        c->rebuild_cap = cap;
        c->P.rebuild_flag = (GMEM int *)c->rebuild_flag;
        HIPCHECK(hipMemcpy(c->d_P, &c->P, sizeof(DP), hipMemcpyHostToDevice));
        if (c->gexec) {
            hipGraphExecDestroy(c->gexec);
            c->gexec = nullptr;
        }
        if (c->graph) {
            hipGraphDestroy(c->graph);
            c->graph = nullptr;
        }
    }
    return 0;
}

// Renumber the particles by applying 'order' (new k = old order[k]) to every
// per-particle array; the positions move to the other ping-pong buffer.
static int apply_order(gmd_ctx *c, const int *order) {
    hipStream_t st = c->stream;
    const int N = c->d.N;
    const int grid = std::max(1, std::min((N + 255) / 256, c->cus * 8));
    hipLaunchKernelGGL(k_permute, dim3(grid), dim3(256), 0, st, (const DP *)c->d_P, c->buf0, order, c->tb);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(c->vel, c->tb.vel, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(c->frc, c->tb.frc, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(c->rref, c->tb.rref, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(c->obeta, c->tb.obeta, (size_t)N * sizeof(float4), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(c->sc, c->tb.sc, (size_t)5 * N * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(c->rim, c->tb.rim, (size_t)3 * N * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipMemcpyAsync(c->perm, c->tb.perm, (size_t)N * sizeof(int), hipMemcpyDeviceToDevice, st));
    c->buf0 ^= 1;
    c->need_rebuild = true;   // list indices refer to the old numbering
    return 0;
}

// cell order of the current positions
static int reorder(gmd_ctx *c) {
    hipStream_t st = c->stream;
    const int N = c->d.N, nc = c->nkeys;
    const int grid = std::max(1, std::min((N + 255) / 256, c->cus * 8));
    const int gridc = std::max(1, std::min((nc + 255) / 256, c->cus * 8));
    HIPCHECK(hipMemsetAsync(c->scnt, 0, (size_t)nc * sizeof(int), st));
    hipLaunchKernelGGL(k_sort_count, dim3(grid), dim3(256), 0, st, (const DP *)c->d_P, c->buf0, c->skey, c->scnt);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, c->scnt, nc);
    HIPCHECK(hipMemcpyAsync(c->scur, c->scnt, (size_t)nc * sizeof(int), hipMemcpyDeviceToDevice, st));
    hipLaunchKernelGGL(k_sort_scatter, dim3(grid), dim3(256), 0, st, N, (const int *)c->skey, c->scur, c->sorder);
    hipLaunchKernelGGL(k_sort_cells, dim3(gridc), dim3(256), 0, st, (const int *)c->scnt, nc, N, c->sorder);
    HIPCHECK(hipGetLastError());
    if (apply_order(c, c->sorder)) return -1;
    c->permuted = true;
    c->since_reorder = 0;
    return 0;
}

// back to the caller's numbering (before host data arrive in that order)
static int unpermute(gmd_ctx *c) {
    if (!c->permuted) return 0;
    const int N = c->d.N;
    const int grid = std::max(1, std::min((N + 255) / 256, c->cus * 8));
    hipLaunchKernelGGL(k_perm_inverse, dim3(grid), dim3(256), 0, c->stream, (const DP *)c->d_P, c->sorder);
    HIPCHECK(hipGetLastError());
    if (apply_order(c, c->sorder)) return -1;
    HIPCHECK(hipStreamSynchronize(c->stream));
    c->permuted = false;
    return 0;
}

// the permutation on the host (internal index -> caller's index), or empty
static int host_perm(gmd_ctx *c, std::vector<int> &perm) {
    perm.clear();
    if (!c->permuted) return 0;
    perm.resize(c->d.N);
    HIPCHECK(hipMemcpy(perm.data(), c->perm, (size_t)c->d.N * sizeof(int), hipMemcpyDeviceToHost));
    return 0;
}

// per-run set-up of one system on its stream: control block, rebuild flags,
// saver rows, the cell grid of the starting positions and the skin test
static int run_prepare(gmd_ctx *c, float t0, int steps) {
    if (steps < 0) return fail("gmd_run: steps must be non-negative");
    if (ensure_flags(c, steps)) return -1;
    if (c->reorder_every > 0 && steps > 0 && c->since_reorder >= c->reorder_every)
        if (reorder(c)) return -1;
    hipStream_t st = c->stream;
    const DP *dP = c->d_P;
    HIPCHECK(hipMemsetAsync(c->rebuild_flag, 0, (size_t)(steps + 2) * sizeof(int), st));
    if (c->need_rebuild) {
        int one = 1;
        HIPCHECK(hipMemcpyAsync(c->rebuild_flag, &one, sizeof(int), hipMemcpyHostToDevice, st));
    }
    if (c->d.scalar_save_every > 0)
        HIPCHECK(hipMemsetAsync(c->saver_out, 0, (size_t)c->d.scalar_saves_per_block * 6 * sizeof(float), st));
    hipLaunchKernelGGL(k_begin_run, dim3(1), dim3(64), 0, st, c->ctrl, steps, t0, c->buf0, c->lstep);
    if (c->d.nblist_kind == GMD_NBLIST_LINKED_LISTS)
        hipLaunchKernelGGL(k_cells_clear, dim3(std::max(1, std::min((3 * c->P.ncells + 255) / 256, c->cus * 8))),
                           dim3(256), 0, st, dP);
    hipLaunchKernelGGL(k_run_setup, dim3(c->grid_aux), dim3(c->block), 0, st, dP);
    hipLaunchKernelGGL(k_setup_finish, dim3(1), dim3(64), 0, st, dP);
    HIPCHECK(hipGetLastError());
    return 0;
}

// after the stream has drained: host-side state of the run and its errors
static int run_complete(gmd_ctx *c, int steps) {
    c->buf0 ^= (steps & 1);
    c->lstep += steps;
    c->since_reorder += steps;
    c->need_rebuild = false;
    Ctrl h;
    HIPCHECK(hipMemcpy(&h, c->ctrl, sizeof h, hipMemcpyDeviceToHost));
    if (h.cell_overflow > 0)
        return fail("linked lists: %d particles fell into one cell, more than the cell capacity %d "
                    "(use 'N squared' or a larger skin)", h.cell_overflow, c->P.ccap);
    return 0;
}

static int run_impl(gmd_ctx *c, float t0, int steps) {
    if (run_prepare(c, t0, steps)) return -1;
    hipStream_t st = c->stream;
    const DP *dP = c->d_P;
    HIPCHECK(hipEventRecord(c->ev0, st));
    if (c->d.gridsync) {
        void *args[] = {(void *)&dP};
        HIPCHECK(hipLaunchCooperativeKernel((const void *)c->kern.pers, dim3(c->grid_pers), dim3(c->block), args,
                                            (unsigned)c->lds_pers, st));
    } else if (c->d.use_graph) {
        if (!c->gexec) {
            HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < GRAPH_STEPS; ++k) {
                hipLaunchKernelGGL(c->kern.step, dim3(c->grid_step), dim3(c->block), c->lds_step, st, dP);
                hipLaunchKernelGGL(k_tail, dim3(1), dim3(256), 0, st, dP, c->d.integrator);
            }
            HIPCHECK(hipStreamEndCapture(st, &c->graph));
            HIPCHECK(hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0));
        }
        const int launches = (steps + 1 + GRAPH_STEPS - 1) / GRAPH_STEPS;
        for (int k = 0; k < launches; ++k) HIPCHECK(hipGraphLaunch(c->gexec, st));
    } else {
        for (int s = 0; s <= steps; ++s) {
            hipLaunchKernelGGL(c->kern.step, dim3(c->grid_step), dim3(c->block), c->lds_step, st, dP);
            hipLaunchKernelGGL(k_tail, dim3(1), dim3(256), 0, st, dP, c->d.integrator);
        }
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipEventRecord(c->ev1, st));
    HIPCHECK(hipEventSynchronize(c->ev1));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1));
    return run_complete(c, steps);
}

int gmd_run(gmd_ctx *c, float time_zero, int steps) {
    if (!c) return fail("gmd_run: null context");
    if (!c->own_stream) return fail("gmd_run: this context is a replica of a batch; use gmd_batch_run");
    if (c->d.integrator == GMD_INTEGRATOR_NONE && steps > 0)
        return fail("gmd_run: context has no integrator; only steps=0 (evaluation) is possible");
    return run_impl(c, time_zero, steps);
}

int gmd_evaluate(gmd_ctx *c) {
    if (!c) return fail("gmd_evaluate: null context");
    if (!c->own_stream) return fail("gmd_evaluate: this context is a replica of a batch");
    return run_impl(c, 0.0f, 0);
}

int gmd_profile_steps(gmd_ctx *c, float time_zero, int steps, float *step_ms, float *tail_ms) {
    if (!c || !step_ms || !tail_ms) return fail("gmd_profile_steps: null argument");
    if (!c->own_stream) return fail("gmd_profile_steps: this context is a replica of a batch");
    if (c->d.gridsync) return fail("gmd_profile_steps: the persistent driver runs every step in one launch");
    if (c->d.integrator == GMD_INTEGRATOR_NONE && steps > 0) return fail("gmd_profile_steps: no integrator");
    if (run_prepare(c, time_zero, steps)) return -1;
    hipStream_t st = c->stream;
    const DP *dP = c->d_P;
    const int n = steps + 1;
    std::vector<hipEvent_t> ev((size_t)3 * n, nullptr);
    int rc = 0;
    for (auto &e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            rc = fail("gmd_profile_steps: hipEventCreate failed");
            break;
        }
    if (rc == 0) {
        for (int s = 0; s < n; ++s) {
            hipEventRecord(ev[3 * s], st);
            hipLaunchKernelGGL(c->kern.step, dim3(c->grid_step), dim3(c->block), c->lds_step, st, dP);
            hipEventRecord(ev[3 * s + 1], st);
            hipLaunchKernelGGL(k_tail, dim3(1), dim3(256), 0, st, dP, c->d.integrator);
            hipEventRecord(ev[3 * s + 2], st);
        }
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
            rc = fail("gmd_profile_steps: launch failed");
    }
    if (rc == 0) {
        double a = 0.0, b = 0.0;
        for (int s = 0; s < n; ++s) {
            float x = 0.f, y = 0.f;
            hipEventElapsedTime(&x, ev[3 * s], ev[3 * s + 1]);
            hipEventElapsedTime(&y, ev[3 * s + 1], ev[3 * s + 2]);
            a += x;
            b += y;
        }
        *step_ms = (float)(a / n);
        *tail_ms = (float)(b / n);
    }
    for (auto &e : ev)
        if (e) hipEventDestroy(e);
    if (rc) return rc;
    return run_complete(c, steps);
}

int gmd_last_run_ms(gmd_ctx *c, float *ms) {
    if (!c || !ms) return fail("gmd_last_run_ms: null argument");
    *ms = c->last_ms;
    return 0;
}

int gmd_read_nbflag(gmd_ctx *c, int nbflag[3]) {
    if (!c || !nbflag) return fail("gmd_read_nbflag: null argument");
    Ctrl h;
    HIPCHECK(hipMemcpy(&h, c->ctrl, sizeof h, hipMemcpyDeviceToHost));
    for (int k = 0; k < 3; ++k) nbflag[k] = h.nbflag[k];
    return 0;
}

int gmd_read_nblist(gmd_ctx *c, int *out) {
    if (!c || !out) return fail("gmd_read_nblist: null argument");
    const int N = c->d.N, M = c->d.max_nbs, tp = c->d.tp;
    std::vector<int> nbl(c->nbl_len), nbc(N);
    HIPCHECK(hipMemcpy(nbl.data(), c->nbl, nbl.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(nbc.data(), c->nbc, (size_t)N * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> perm;
    if (host_perm(c, perm)) return -1;
    auto at = [&](int i) { return perm.empty() ? i : perm[i]; };
    for (int i = 0; i < N; ++i) {
        int *row = out + (size_t)at(i) * (M + 1);
        int n = std::min(nbc[i], M);
        for (int k = 0; k < n; ++k) row[k] = at(nbl[((size_t)(k / tp) * N + i) * tp + (k % tp)]);
        for (int k = n; k < M; ++k) row[k] = 0;
        row[M] = nbc[i];
    }
    return 0;
}

int gmd_set_reorder_interval(gmd_ctx *c, int steps) {
    if (!c) return fail("gmd_set_reorder_interval: null context");
    if (steps < 0) return fail("gmd_set_reorder_interval: steps must be non-negative");
    if (steps > 0 && !c->sorder)
        return fail("gmd_set_reorder_interval: renumbering needs linked lists and no exclusions");
    c->reorder_every = steps;
    return 0;
}

int gmd_force_rebuild(gmd_ctx *c) {
    if (!c) return fail("gmd_force_rebuild: null context");
    c->need_rebuild = true;
    return 0;
}

int gmd_get_thermostat(gmd_ctx *c, float state[2]) {
    if (!c || !state) return fail("gmd_get_thermostat: null argument");
    Ctrl h;
    HIPCHECK(hipMemcpy(&h, c->ctrl, sizeof h, hipMemcpyDeviceToHost));
    state[0] = h.xi;
    state[1] = h.ksum;
    return 0;
}

int gmd_set_thermostat(gmd_ctx *c, const float state[2]) {
    if (!c || !state) return fail("gmd_set_thermostat: null argument");
    HIPCHECK(hipMemcpy(&c->ctrl->xi, state, 2 * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int gmd_read_scalar_output(gmd_ctx *c, float *out) {
    if (!c || !out) return fail("gmd_read_scalar_output: null argument");
    if (c->d.scalar_save_every <= 0) return fail("gmd_read_scalar_output: no scalar saver configured");
    HIPCHECK(hipMemcpy(out, c->saver_out, (size_t)c->d.scalar_saves_per_block * 6 * sizeof(float),
                       hipMemcpyDeviceToHost));
    return 0;
}

// ---------------------------------------------------------------------------
// RDF
// ---------------------------------------------------------------------------
static int rdf_bin_host(float x, double half, double bw, int bins) {
    // calculator_radial_distribution.py:101-104 with numba's float64 promotion
    if (!((double)x < half * half)) return bins;
    float dd = std::sqrt(x);
    if (!((double)dd < half)) return bins;
    int b = (int)std::floor((double)dd / bw);
    return std::min(b, bins);
}

static void rdf_thresholds(const float box[3], int bins, std::vector<float> &thr) {
    float mn = std::min(box[0], std::min(box[1], box[2]));
    double half = (double)mn / 2.0;
    double bw = half / bins;
    thr.assign(bins + 1, 0.f);
    for (int k = 1; k <= bins; ++k) {
        // smallest non-negative float x with bin(x) >= k (bin is monotone in x)
        uint32_t lo = 0, hi = 0x7f800000u;  // +inf
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            float x;
            memcpy(&x, &mid, 4);
            if (rdf_bin_host(x, half, bw, bins) >= k)
                hi = mid;
            else
                lo = mid + 1;
        }
        float x;
        memcpy(&x, &lo, 4);
        thr[k] = x;
    }
}

static size_t rdf_lds_bytes(int nclass, int bins, int nt) {
    const size_t nh = (size_t)nclass * bins;
    return ((nh * 4 + 15) & ~(size_t)15) + (size_t)bins * 8 + 256 * 16 + (size_t)nt * nt * 4;
}

int gmd_rdf_create(int N, int ntypes, int bins, gmd_rdf **out) {
    if (!out) return fail("gmd_rdf_create: null output");
    *out = nullptr;
    if (N < 1 || ntypes < 1 || ntypes > MAX_TYPES || bins < 1) return fail("gmd_rdf_create: bad sizes");
    gmd_rdf *r = new gmd_rdf();
    r->N = N;
    r->nt = ntypes;
    r->bins = bins;
    r->nclass = ntypes * (ntypes + 1) / 2;
    const size_t nh = (size_t)r->nclass * bins;
    if (rdf_lds_bytes(r->nclass, bins, ntypes) > 64 * 1024) {
        delete r;
        return fail("gmd_rdf_create: %d type pairs x %d bins do not fit the LDS histogram", ntypes * (ntypes + 1) / 2,
                    bins);
    }
    int dev = 0, cus = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const long long T = (N + 255) / 256;
    const long long npairs = T * (T + 1) / 2;
    r->grid = (int)std::max(1LL, std::min(npairs, (long long)cus * 4));
    if (hipMalloc(&r->pos, (size_t)N * sizeof(float4)) != hipSuccess ||
        hipMalloc(&r->tbox, (size_t)2 * T * sizeof(float4)) != hipSuccess ||
        hipMalloc(&r->ptype, (size_t)N * sizeof(int)) != hipSuccess ||
        hipMalloc(&r->d_thr, (size_t)bins * sizeof(float2)) != hipSuccess ||
        hipMalloc(&r->part, (size_t)r->grid * nh * sizeof(unsigned)) != hipSuccess ||
        hipMalloc(&r->d_out, nh * sizeof(unsigned long long)) != hipSuccess) {
        gmd_rdf_destroy(r);
        return fail("gmd_rdf_create: out of device memory");
    }
    if (hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking) != hipSuccess) {
        gmd_rdf_destroy(r);
        return fail("gmd_rdf_create: stream creation failed");
    }
    *out = r;
    return 0;
}

int gmd_rdf_destroy(gmd_rdf *r) {
    if (!r) return 0;
    void *ptrs[] = {r->pos, r->tbox, r->ptype, r->d_thr, r->part, r->d_out};
    for (void *p : ptrs)
        if (p) hipFree(p);
    if (r->stream) hipStreamDestroy(r->stream);
    delete r;
    return 0;
}

static int rdf_run(gmd_rdf *r, const float4 *pos, const int *d_ptype, const float box[3], unsigned long long *counts) {
    if (box[0] != r->box[0] || box[1] != r->box[1] || box[2] != r->box[2] || r->thr.empty()) {
        rdf_thresholds(box, r->bins, r->thr);
        std::vector<float2> t2(r->bins);
        for (int k = 0; k < r->bins; ++k) t2[k] = make_float2(r->thr[k], r->thr[k + 1]);
        HIPCHECK(hipMemcpy(r->d_thr, t2.data(), t2.size() * sizeof(float2), hipMemcpyHostToDevice));
        memcpy(r->box, box, sizeof r->box);
    }
    RdfP R;
    R.N = r->N;
    R.nt = r->nt;
    R.bins = r->bins;
    R.nclass = r->nclass;
    R.tile = 256;
    R.ntiles = (r->N + 255) / 256;
    R.npairs = (long long)R.ntiles * (R.ntiles + 1) / 2;
    R.L0 = box[0];
    R.L1 = box[1];
    R.L2 = box[2];
    {
        float mn = std::min(box[0], std::min(box[1], box[2]));
        double bw = ((double)mn / 2.0) / r->bins;
        R.inv_bw = (float)(1.0 / bw);
    }
    R.pos = (cpos_t)pos;
    R.ptype = (const GMEM int *)d_ptype;
    R.thr2 = (const GMEM float2 *)r->d_thr;
    R.out_thr = r->thr[r->bins];
    R.part = (GMEM unsigned *)r->part;
    R.box = (const GMEM float4 *)r->tbox;
    R.skip2 = R.out_thr;
    const size_t nh = (size_t)r->nclass * r->bins;
    HIPCHECK(hipMemsetAsync(r->d_out, 0, nh * sizeof(unsigned long long), r->stream));
    hipLaunchKernelGGL(k_rdf_boxes, dim3(R.ntiles), dim3(256), 0, r->stream, pos, r->N, box[0], box[1], box[2],
                       r->tbox);
    hipLaunchKernelGGL(k_rdf, dim3(r->grid), dim3(256), rdf_lds_bytes(r->nclass, r->bins, r->nt), r->stream, R);
    HIPCHECK(hipGetLastError());
    const int per = 16;
    hipLaunchKernelGGL(k_rdf_reduce, dim3((unsigned)((nh + 255) / 256), (unsigned)((r->grid + per - 1) / per)),
                       dim3(256), 0, r->stream, (const unsigned *)r->part, r->grid, (int)nh, per, r->d_out);
    HIPCHECK(hipGetLastError());
    std::vector<unsigned long long> h(nh);
    HIPCHECK(hipMemcpyAsync(h.data(), r->d_out, nh * sizeof(unsigned long long), hipMemcpyDeviceToHost, r->stream));
    HIPCHECK(hipStreamSynchronize(r->stream));
    const int nt = r->nt, bins = r->bins;
    for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
            int lo = std::min(a, b), hi = std::max(a, b);
            int cls = lo * nt - (lo * (lo - 1)) / 2 + (hi - lo);
            for (int k = 0; k < bins; ++k) {
                unsigned long long v = h[(size_t)cls * bins + k];
                counts[((size_t)a * nt + b) * bins + k] = (a == b) ? 2 * v : v;
            }
        }
    return 0;
}

int gmd_rdf_compute_ctx(gmd_rdf *r, gmd_ctx *c, const int *ptype, unsigned long long *counts) {
    if (!r || !c || !counts) return fail("gmd_rdf_compute_ctx: null argument");
    if (c->d.N != r->N) return fail("gmd_rdf_compute_ctx: N mismatch (%d vs %d)", c->d.N, r->N);
    HIPCHECK(hipStreamSynchronize(c->stream));
    const int *dp = nullptr;
    if (ptype) {
        for (int i = 0; i < r->N; ++i)
            if (ptype[i] < 0 || ptype[i] >= r->nt) return fail("gmd_rdf_compute_ctx: type out of range");
        std::vector<int> perm, pt(ptype, ptype + r->N);
        if (host_perm(c, perm)) return -1;
        if (!perm.empty())
            for (int i = 0; i < r->N; ++i) pt[i] = ptype[perm[i]];   // types in the context's numbering
        HIPCHECK(hipMemcpy(r->ptype, pt.data(), (size_t)r->N * sizeof(int), hipMemcpyHostToDevice));
        dp = r->ptype;
    } else if (c->d.ntypes > r->nt) {
        return fail("gmd_rdf_compute_ctx: context has %d types, calculator %d", c->d.ntypes, r->nt);
    }
    return rdf_run(r, c->buf0 ? c->pos1 : c->pos0, dp, c->d.box, counts);
}

int gmd_rdf_compute_host(gmd_rdf *r, const float *pos, const int *ptype, const float box[3],
                         unsigned long long *counts) {
    if (!r || !pos || !ptype || !box || !counts) return fail("gmd_rdf_compute_host: null argument");
    for (int i = 0; i < r->N; ++i)
        if (ptype[i] < 0 || ptype[i] >= r->nt) return fail("gmd_rdf_compute_host: type out of range");
    std::vector<float4> h(r->N);
    for (int i = 0; i < r->N; ++i) h[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], i2f(ptype[i]));
    HIPCHECK(hipMemcpy(r->pos, h.data(), h.size() * sizeof(float4), hipMemcpyHostToDevice));
    return rdf_run(r, r->pos, nullptr, box, counts);
}


// ============================================================================
// Batches of independent replicas
// ============================================================================
struct gmd_batch {
    std::vector<gmd_ctx *> reps;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int driver = GMD_BATCH_PERSISTENT, integrator = 0, block = 256, grid = 0, cus = 0;
    size_t lds = 0;
    Kernels kern = {nullptr, nullptr, nullptr, nullptr};
    std::vector<BatchBlk> h_blk;
    BatchBlk *d_blk = nullptr;
    const DP **d_Ps = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    float last_ms = 0.f;
};

static int batch_free(gmd_batch *b) {
    if (!b) return 0;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    for (gmd_ctx *c : b->reps) ctx_free(c);
    if (b->d_blk) hipFree(b->d_blk);
    if (b->d_Ps) hipFree(b->d_Ps);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

// Grid layout of a batch.  Consecutive workgroups of a launch go to the
// device's 8 XCDs in turn, so grid slot k*8 + x runs on XCD x.  The blocks
// are packed into 8 bins of T = ceil(total/8) slots, largest replica first,
// each into the fullest bin that still holds it whole (its positions and
// lists then live in one XCD's L2); a replica that fits no bin whole is
// spread over the emptiest bins.  At most 7 slots stay idle (rep = -1).
static void batch_layout(const std::vector<int> &nb, std::vector<BatchBlk> &out) {
    const int NX = 8;
    long long total = 0;
    for (int n : nb) total += n;
    const long long T = (total + NX - 1) / NX;
    std::vector<int> order(nb.size());
    for (size_t r = 0; r < nb.size(); ++r) order[r] = (int)r;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return nb[a] > nb[b2]; });
    std::vector<std::vector<BatchBlk>> bins(NX);
    for (int r : order) {
        int best = -1;
        for (int x = 0; x < NX; ++x) {
            const long long room = T - (long long)bins[x].size();
            if (room >= nb[r] && (best < 0 || room < T - (long long)bins[best].size())) best = x;
        }
        if (best >= 0) {
            for (int k = 0; k < nb[r]; ++k) bins[best].push_back(BatchBlk{r, k, nb[r], 0});
            continue;
        }
        int k = 0;
        while (k < nb[r]) {
            int x = 0;
            for (int y = 1; y < NX; ++y)
                if (bins[y].size() < bins[x].size()) x = y;
            while (k < nb[r] && (long long)bins[x].size() < T) bins[x].push_back(BatchBlk{r, k++, nb[r], 0});
        }
    }
    out.assign((size_t)T * NX, BatchBlk{-1, 0, 0, 0});
    for (int x = 0; x < NX; ++x)
        for (size_t k = 0; k < bins[x].size(); ++k) out[k * NX + x] = bins[x][k];
}

int gmd_batch_create(const gmd_desc *descs, int nrep, const int *blocks, int driver, gmd_batch **out) {
    if (!descs || !out) return fail("gmd_batch_create: null argument");
    *out = nullptr;
    if (nrep < 1) return fail("gmd_batch_create: need at least one replica");
    if (driver != GMD_BATCH_PERSISTENT && driver != GMD_BATCH_GRAPH)
        return fail("gmd_batch_create: unknown driver %d", driver);
    const gmd_desc &d0 = descs[0];
    for (int r = 1; r < nrep; ++r) {
        const gmd_desc &d = descs[r];
        if (d.pair_base != d0.pair_base || d.cutoff != d0.cutoff || d.integrator != d0.integrator ||
            d.ntypes != d0.ntypes || d.tp != d0.tp || d.pb != d0.pb)
            return fail("gmd_batch_create: replica %d differs from replica 0 in pair function, cutoff, integrator, "
                        "number of types, tp or pb (one kernel runs every replica)", r);
    }
    if (d0.integrator == GMD_INTEGRATOR_NONE) return fail("gmd_batch_create: replicas need an integrator");
    gmd_batch *b = new gmd_batch();
    b->driver = driver;
    b->integrator = d0.integrator;
#define BTRY(x)                                                             \
    do {                                                                    \
        hipError_t e_ = (x);                                                \
        if (e_ != hipSuccess) {                                             \
            batch_free(b);                                                  \
            return fail("gmd_batch_create: %s: %s", #x, hipGetErrorString(e_)); \
        }                                                                   \
    } while (0)
    BTRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    BTRY(hipEventCreate(&b->ev0));
    BTRY(hipEventCreate(&b->ev1));
    std::vector<int> nb(nrep);
    for (int r = 0; r < nrep; ++r) {
        gmd_desc d = descs[r];
        d.gridsync = 0;
        d.use_graph = 0;
        gmd_ctx *c = nullptr;
        if (ctx_create(&d, &c, b->stream)) {
            std::string msg = g_err;
            batch_free(b);
            return fail("gmd_batch_create: replica %d: %s", r, msg.c_str());
        }
        b->reps.push_back(c);
        nb[r] = blocks ? blocks[r] : c->grid_step;
        const int cap = std::max(c->grid_step, c->cus * 32);
        if (nb[r] < 1 || nb[r] > cap) {
            batch_free(b);
            return fail("gmd_batch_create: replica %d: %d blocks requested, allowed 1..%d", r, nb[r], cap);
        }
        c->P.step_grid = nb[r];
        BTRY(hipMemcpy(c->d_P, &c->P, sizeof(DP), hipMemcpyHostToDevice));
        b->lds = std::max(b->lds, c->lds_step);
        b->block = c->block;
        b->cus = c->cus;
    }
    b->kern = b->reps[0]->kern;
    batch_layout(nb, b->h_blk);
    b->grid = (int)b->h_blk.size();
    if (driver == GMD_BATCH_PERSISTENT) {
        int per_cu = 0;
        BTRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)b->kern.bpers, b->block, b->lds));
        if (per_cu * b->cus < b->grid) {
            const int g = b->grid, room = per_cu * b->cus, bs = b->block;
            batch_free(b);
            return fail("gmd_batch_create: the persistent batch needs %d co-resident blocks of %d threads, the device "
                        "holds %d; assign fewer blocks per replica", g, bs, room);
        }
    }
    BTRY(hipMalloc((void **)&b->d_blk, b->h_blk.size() * sizeof(BatchBlk)));
    BTRY(hipMemcpy(b->d_blk, b->h_blk.data(), b->h_blk.size() * sizeof(BatchBlk), hipMemcpyHostToDevice));
    std::vector<const DP *> ps(nrep);
    for (int r = 0; r < nrep; ++r) ps[r] = b->reps[r]->d_P;
    BTRY(hipMalloc((void **)&b->d_Ps, nrep * sizeof(DP *)));
    BTRY(hipMemcpy(b->d_Ps, ps.data(), nrep * sizeof(DP *), hipMemcpyHostToDevice));
#undef BTRY
    *out = b;
    return 0;
}

int gmd_batch_destroy(gmd_batch *b) { return batch_free(b); }

int gmd_batch_size(gmd_batch *b, int *nrep, int *grid) {
    if (!b) return fail("gmd_batch_size: null batch");
    if (nrep) *nrep = (int)b->reps.size();
    if (grid) *grid = b->grid;
    return 0;
}

int gmd_batch_context(gmd_batch *b, int rep, gmd_ctx **out) {
    if (!b || !out) return fail("gmd_batch_context: null argument");
    if (rep < 0 || rep >= (int)b->reps.size()) return fail("gmd_batch_context: replica %d out of range", rep);
    *out = b->reps[rep];
    return 0;
}

int gmd_batch_run(gmd_batch *b, float time_zero, int steps) {
    if (!b) return fail("gmd_batch_run: null batch");
    for (size_t r = 0; r < b->reps.size(); ++r)
        if (run_prepare(b->reps[r], time_zero, steps)) {
            std::string msg = g_err;
            return fail("gmd_batch_run: replica %d: %s", (int)r, msg.c_str());
        }
    // the rebuild-flag arrays may have been reallocated by the preparation
    hipStream_t st = b->stream;
    HIPCHECK(hipEventRecord(b->ev0, st));
    const DP *const *Ps = b->d_Ps;
    const BatchBlk *blk = b->d_blk;
    if (b->driver == GMD_BATCH_PERSISTENT) {
        void *args[] = {(void *)&Ps, (void *)&blk};
        HIPCHECK(hipLaunchCooperativeKernel((const void *)b->kern.bpers, dim3(b->grid), dim3(b->block), args,
                                            (unsigned)b->lds, st));
    } else {
        if (!b->gexec) {
            HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < GRAPH_STEPS; ++k) {
                hipLaunchKernelGGL(b->kern.bstep, dim3(b->grid), dim3(b->block), b->lds, st, Ps, blk);
                hipLaunchKernelGGL(k_batch_tail, dim3((unsigned)b->reps.size()), dim3(256), 0, st, Ps, b->integrator);
            }
            HIPCHECK(hipStreamEndCapture(st, &b->graph));
            HIPCHECK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        const int launches = (steps + 1 + GRAPH_STEPS - 1) / GRAPH_STEPS;
        for (int k = 0; k < launches; ++k) HIPCHECK(hipGraphLaunch(b->gexec, st));
    }
    HIPCHECK(hipEventRecord(b->ev1, st));
    HIPCHECK(hipEventSynchronize(b->ev1));
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
    for (size_t r = 0; r < b->reps.size(); ++r)
        if (run_complete(b->reps[r], steps)) {
            std::string msg = g_err;
            return fail("gmd_batch_run: replica %d: %s", (int)r, msg.c_str());
        }
    return 0;
}

int gmd_batch_last_run_ms(gmd_batch *b, float *ms) {
    if (!b || !ms) return fail("gmd_batch_last_run_ms: null argument");
    *ms = b->last_ms;
    return 0;
}

}  // extern "C"

#include "gmd_adjoint.h"
