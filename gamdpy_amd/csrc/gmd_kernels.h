// gmd_kernels.h -- the per-timestep phases as device functions.
//
// Each phase is written once and called from two drivers:
//   * k_persistent: one cooperative launch per timeblock; phases separated by
//     a software grid barrier (Simulation.py:306-328 grid.sync() loop);
//   * kernel-per-phase: one small __global__ wrapper per phase, replayed from
//     a captured hipGraph (Simulation.py:330-346 host loop).
#pragma once
#include "gmd_device.h"

// neighbour-list slot of entry k of particle i (interleaved by tp lanes)
__device__ __forceinline__ size_t nbl_index(const DP &P, int i, int k) {
    return ((size_t)(k >> P.tp_shift) * (size_t)P.N + (size_t)i) * (size_t)P.tp + (size_t)(k & (P.tp - 1));
}

__device__ __forceinline__ const float4 *cur_pos(const DP &P, int s) {
    return ((P.ctrl->buf0 ^ (s & 1)) == 0) ? P.pos0 : P.pos1;
}
__device__ __forceinline__ float4 *next_pos(const DP &P, int s) {
    return ((P.ctrl->buf0 ^ (s & 1)) == 0) ? P.pos1 : P.pos0;
}

__device__ __forceinline__ int ld_flag(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// nblist.py:112-116 (NIII half list) or j != i
__device__ __forceinline__ bool pair_accept(const DP &P, int i, int j) {
    if (P.niii) {
        int two_dn = 2 * (j - i);
        return (two_dn > 0 && two_dn <= P.N) || (two_dn < -P.N);
    }
    return j != i;
}

__device__ __forceinline__ bool excluded(const DP &P, int i, int j) {
    if (P.excl == nullptr) return false;
    const int *row = P.excl + (size_t)i * P.excl_cols;
    int n = row[P.excl_cols - 1];
    for (int k = 0; k < n; ++k)
        if (row[k] == j) return true;
    return false;
}

__device__ inline void load_pair_table(const DP &P, PairP *pp) {
    int nt2 = P.ntypes * P.ntypes;
    for (int t = threadIdx.x; t < nt2; t += blockDim.x) {
        const float *q = P.raw_params + 3 * t;
        make_pairp(P.pair_base, P.cutoff, q[0], q[1], q[2], pp[t]);
    }
}

// ---------------------------------------------------------------------------
// Linked-list rebuild (nblist_linked_lists.py:117-218), as a counting sort.
// R1: cell of each particle + rank inside the cell
// ---------------------------------------------------------------------------
__device__ inline void ll_assign(const DP &P, const float4 *pos, int gtid, int gstride) {
    const double L0 = P.box0, L1 = P.box1, L2 = P.box2;
    for (int i = gtid; i < P.N; i += gstride) {
        float4 r = pos[i];
        // nblist_linked_lists.py:127, evaluated in float64 as numba does
        long long cx = (long long)floor(((double)r.x + 0.5 * L0) * (double)P.ncx / L0);
        long long cy = (long long)floor(((double)r.y + 0.5 * L1) * (double)P.ncy / L1);
        long long cz = (long long)floor(((double)r.z + 0.5 * L2) * (double)P.ncz / L2);
        cx %= P.ncx; if (cx < 0) cx += P.ncx;
        cy %= P.ncy; if (cy < 0) cy += P.ncy;
        cz %= P.ncz; if (cz < 0) cz += P.ncz;
        int c = (int)((cx * P.ncy + cy) * P.ncz + cz);
        P.cell_of[i] = c;
        P.cell_rank[i] = atomicAdd(&P.cell_count[c], 1);
    }
}

// R2: exclusive scan of cell counts -> cell_start[0..ncells] (one block)
__device__ inline void ll_scan_block(const DP &P, int *lds_tmp /* blockDim ints */) {
    const int nt = blockDim.x, t = threadIdx.x;
    const int n = P.ncells;
    const int per = (n + nt - 1) / nt;
    const int b = t * per, e = min(n, b + per);
    int sum = 0;
    for (int c = b; c < e; ++c) sum += P.cell_count[c];
    lds_tmp[t] = sum;
    __syncthreads();
    // Hillis-Steele inclusive scan over thread sums
    for (int off = 1; off < nt; off <<= 1) {
        int v = (t >= off) ? lds_tmp[t - off] : 0;
        __syncthreads();
        lds_tmp[t] += v;
        __syncthreads();
    }
    int run = lds_tmp[t] - sum;
    for (int c = b; c < e; ++c) {
        P.cell_start[c] = run;
        run += P.cell_count[c];
    }
    if (t == nt - 1) P.cell_start[n] = lds_tmp[t];
    __syncthreads();
}

// R3: scatter particle indices into cell order
__device__ inline void ll_scatter(const DP &P, int gtid, int gstride) {
    for (int i = gtid; i < P.N; i += gstride) P.cell_members[P.cell_start[P.cell_of[i]] + P.cell_rank[i]] = i;
}

// R4: sort members of each cell by index (deterministic list order) and clear
// the counts for the next rebuild
__device__ inline void ll_sort_cells(const DP &P, int gtid, int gstride) {
    for (int c = gtid; c < P.ncells; c += gstride) {
        int b = P.cell_start[c], e = P.cell_start[c + 1];
        for (int x = b + 1; x < e; ++x) {
            int v = P.cell_members[x];
            int y = x - 1;
            while (y >= b && P.cell_members[y] > v) {
                P.cell_members[y + 1] = P.cell_members[y];
                --y;
            }
            P.cell_members[y + 1] = v;
        }
        P.cell_count[c] = 0;
    }
}

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// R5: one wavefront per cell.  The 125 stencil cells (5x5x5, cell >= (rc+skin)/2,
// nblist_linked_lists.py:151-182) are staged into LDS in chunks of llcap
// candidates; every particle of the cell tests them 64 at a time and appends
// accepted neighbours with a ballot, so the list order is deterministic
// (stencil order, then particle index).
__device__ inline void ll_build(const DP &P, const float4 *pos, char *lds, int wave_global, int wave_stride) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int llcap = P.llcap;
    float4 *cand = (float4 *)lds + (size_t)wid * llcap;
    const float L0 = P.box0, L1 = P.box1, L2 = P.box2;
    const unsigned long long lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int c = wave_global; c < P.ncells; c += wave_stride) {
        const int cst = P.cell_start[c];
        const int ccnt = P.cell_start[c + 1] - cst;
        if (ccnt == 0) continue;
        const int cz = c % P.ncz, cy = (c / P.ncz) % P.ncy, cx = c / (P.ncz * P.ncy);
        int st[2], n[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int sc = lane + 64 * h;
            st[h] = 0;
            n[h] = 0;
            if (sc < 125) {
                int ix = sc / 25 - 2, iy = (sc / 5) % 5 - 2, iz = sc % 5 - 2;
                int ox = (cx + ix) % P.ncx; if (ox < 0) ox += P.ncx;
                int oy = (cy + iy) % P.ncy; if (oy < 0) oy += P.ncy;
                int oz = (cz + iz) % P.ncz; if (oz < 0) oz += P.ncz;
                int nc = (ox * P.ncy + oy) * P.ncz + oz;
                st[h] = P.cell_start[nc];
                n[h] = P.cell_start[nc + 1] - st[h];
            }
        }
        // exclusive offsets in stencil order (sc = 0..124)
        int off[2];
        int total = 0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int v = n[h];
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                int t = __shfl_up(v, o, 64);
                if (lane >= o) v += t;
            }
            off[h] = total + v - n[h];
            total += __shfl(v, 63, 64);
        }
        const int M = total;
        for (int pg = 0; pg < ccnt; pg += 64) {
            const int np = min(64, ccnt - pg);
            const int myi = (lane < np) ? P.cell_members[cst + pg + lane] : -1;
            float4 myr = make_float4(0.f, 0.f, 0.f, 0.f);
            if (myi >= 0) myr = pos[myi];
            int mycount = 0;
            for (int cb = 0; cb < M; cb += llcap) {
                wave_sync_lds();
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    for (int m = 0; m < n[h]; ++m) {
                        int k = off[h] + m - cb;
                        if (k >= 0 && k < llcap) {
                            int j = P.cell_members[st[h] + m];
                            float4 rj = pos[j];
                            rj.w = __int_as_float(j);
                            cand[k] = rj;
                        }
                    }
                }
                wave_sync_lds();
                const int ce = min(llcap, M - cb);
                for (int q = 0; q < np; ++q) {
                    const int iq = __shfl(myi, q, 64);
                    const float rx = __shfl(myr.x, q, 64), ry = __shfl(myr.y, q, 64), rz = __shfl(myr.z, q, 64);
                    int cntq = __shfl(mycount, q, 64);
                    for (int kk = 0; kk < ce; kk += 64) {
                        bool ok = false;
                        int j = -1;
                        if (kk + lane < ce) {
                            float4 cj = cand[kk + lane];
                            j = __float_as_int(cj.w);
                            if (pair_accept(P, iq, j)) {
                                float dx, dy, dz;
                                float d2 = dist_sq_mic(cj.x, cj.y, cj.z, rx, ry, rz, L0, L1, L2, dx, dy, dz);
                                ok = d2 < P.cps2 && !excluded(P, iq, j);
                            }
                        }
                        unsigned long long mask = __ballot(ok);
                        if (ok) {
                            int slot = cntq + __popcll(mask & lt_mask);
                            if (slot < P.max_nbs) P.nbl[nbl_index(P, iq, slot)] = j;
                        }
                        cntq += __popcll(mask);
                    }
                    if (lane == q) mycount = cntq;
                }
            }
            if (myi >= 0) {
                P.nbc[myi] = mycount;
                P.rref[myi] = myr;
                if (mycount > P.max_nbs) atomicMax(&P.ctrl->nbflag[1], mycount);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// N-squared rebuild (nblist.py:89-143): positions staged through LDS tiles,
// one thread per particle, appends in ascending j.  Every thread of every
// block runs the same number of rounds (the tile loads need __syncthreads).
// ---------------------------------------------------------------------------
__device__ inline void nsq_build(const DP &P, const float4 *pos, float4 *tile) {
    const int bd = blockDim.x;
    const int per_round = gridDim.x * bd;
    const int rounds = (P.N + per_round - 1) / per_round;
    const float L0 = P.box0, L1 = P.box1, L2 = P.box2;
    for (int rd = 0; rd < rounds; ++rd) {
        const int i = (rd * gridDim.x + blockIdx.x) * bd + threadIdx.x;
        const bool valid = i < P.N;
        float4 ri = valid ? pos[i] : make_float4(0.f, 0.f, 0.f, 0.f);
        int cnt = 0;
        for (int t0 = 0; t0 < P.N; t0 += bd) {
            __syncthreads();
            int jl = t0 + threadIdx.x;
            if (jl < P.N) tile[threadIdx.x] = pos[jl];
            __syncthreads();
            const int tn = min(bd, P.N - t0);
            if (valid) {
                for (int jj = 0; jj < tn; ++jj) {
                    const int j = t0 + jj;
                    if (!pair_accept(P, i, j)) continue;
                    float4 rj = tile[jj];
                    float dx, dy, dz;
                    float d2 = dist_sq_mic(rj.x, rj.y, rj.z, ri.x, ri.y, ri.z, L0, L1, L2, dx, dy, dz);
                    if (d2 < P.cps2 && !excluded(P, i, j)) {
                        if (cnt < P.max_nbs) P.nbl[nbl_index(P, i, cnt)] = j;
                        ++cnt;
                    }
                }
            }
        }
        if (valid) {
            P.nbc[i] = cnt;
            P.rref[i] = ri;
            if (cnt > P.max_nbs) atomicMax(&P.ctrl->nbflag[1], cnt);
        }
    }
}

// ---------------------------------------------------------------------------
// The fused step: interactions (pair_potential.py:185-246, full-list
// calculator :162-178) for every particle, then -- unless this is the final
// interactions-only evaluation of the block -- the integrator update of the
// same particle (NVE.py:59-105 / NVT.py:188-224 / NVT_Langevin.py:342-385),
// the PBC wrap and the skin test of the next step.  Positions are read from
// pos[cur] and written to pos[next], so no other particle can observe a
// half-updated configuration.  Each thread then holds its contribution to
// the per-step sums (K, U, W, lapU, Fsq, momentum, mass).
// ---------------------------------------------------------------------------
struct StepInfo {
    int s;
    bool integrate;
    bool need_ksum, need_mom, need_save;
    bool apply_shift;
    float sh0, sh1, sh2;
    float xi;
    float lang_amp;
};

__device__ inline StepInfo step_info(const DP &P, int s, int integrator) {
    StepInfo I;
    Ctrl *C = P.ctrl;
    I.s = s;
    I.integrate = (s < C->nsteps) && integrator != 0;
    I.need_ksum = I.integrate && integrator == 2;
    I.need_mom = I.integrate && P.mom_every > 0 && (s % P.mom_every) == 0;
    I.need_save = I.integrate && P.save_every > 0 && (s % P.save_every) == 0 && (s / P.save_every) < P.saves_per_block;
    I.apply_shift = (C->shift_step == s);
    I.sh0 = C->shift[0];
    I.sh1 = C->shift[1];
    I.sh2 = C->shift[2];
    I.xi = C->xi;
    I.lang_amp = 0.f;
    if (integrator == 3) {
        double t = (double)C->t0 + (double)s * (double)P.dt;
        float T = (float)target_temperature(P, t);
        {
#pragma clang fp contract(off)
            I.lang_amp = sqrtf(((2.0f * P.alpha) * T) * P.dt);
        }
    }
    return I;
}

__device__ __forceinline__ void wrap_pbc(float &r, int &dim_img, float L) {
#pragma clang fp contract(off)
    if (r * 2.0f > L) {
        r = r - L;
        dim_img += 1;
    }
    if (r * 2.0f < -L) {
        r = r + L;
        dim_img -= 1;
    }
}

template <int BASE, int CUT, int INTEG>
__device__ inline void phase_step(const DP &P, const StepInfo &I, const PairP *pp, float acc[GMD_NRED]) {
    const int tp = P.tp;
    const int t = threadIdx.x & (tp - 1);
    const int slot = threadIdx.x >> P.tp_shift;
    const int slots = blockDim.x >> P.tp_shift;
    const int s = I.s;
    const float4 *__restrict__ pos = cur_pos(P, s);
    float4 *__restrict__ posn = next_pos(P, s);
    const float L0 = P.box0, L1 = P.box1, L2 = P.box2;
    const int nt = P.ntypes;
    const int N = P.N;
    const float vf = 0.5f / 3.0f;

    for (int i = blockIdx.x * slots + slot; i < N; i += gridDim.x * slots) {
        const float4 ri = pos[i];
        const int ti = __float_as_int(ri.w);
        const PairP *ppi = pp + ti * nt;
        const int cnt = min(P.nbc[i], P.max_nbs);
        float fx = 0.f, fy = 0.f, fz = 0.f, uu = 0.f, ww = 0.f, ll = 0.f;
        const int *nbrow = P.nbl + (size_t)i * tp + t;
        const size_t row_stride = (size_t)N * tp;
        int k = t;
        size_t o = 0;
        for (; k < cnt; k += tp, o += row_stride) {
            const int j = nbrow[o];
            const float4 rj = pos[j];
            float dx, dy, dz;
            const float d2 = dist_sq_mic(rj.x, rj.y, rj.z, ri.x, ri.y, ri.z, L0, L1, L2, dx, dy, dz);
            const PairP &q = ppi[__float_as_int(rj.w)];
            if (d2 < q.cut2) {
                float u, sv, umm;
                pair_fast<BASE, CUT>(d2, q, u, sv, umm);
                fx -= dx * sv;
                fy -= dy * sv;
                fz -= dz * sv;
                ww += d2 * sv;
                uu += u;
                ll += umm - 2.0f * sv;  // (1 - D) s + u'' with D = 3
            }
        }
        // reduce over the tp lanes of this particle
        for (int m = 1; m < tp; m <<= 1) {
            fx += __shfl_xor(fx, m, 64);
            fy += __shfl_xor(fy, m, 64);
            fz += __shfl_xor(fz, m, 64);
            uu += __shfl_xor(uu, m, 64);
            ww += __shfl_xor(ww, m, 64);
            ll += __shfl_xor(ll, m, 64);
        }
        uu *= 0.5f;
        ww *= vf;
        const bool writer = (t == 0);
        float4 v4 = P.vel[i];
        if (I.apply_shift) {
#pragma clang fp contract(off)
            v4.x = v4.x - I.sh0;
            v4.y = v4.y - I.sh1;
            v4.z = v4.z - I.sh2;
        }
        if (writer) {
            P.frc[i] = make_float4(fx, fy, fz, 0.f);
            if (P.compute_u) P.sc[(size_t)SC_U * N + i] = uu;
            if (P.compute_w) P.sc[(size_t)SC_W * N + i] = ww;
            if (P.compute_lap) P.sc[(size_t)SC_LAP * N + i] = ll;
            acc[RED_U] += uu;
            acc[RED_W] += ww;
            acc[RED_LAP] += ll;
        }
        if (!I.integrate) {
            if (writer && I.apply_shift) P.vel[i] = v4;
            continue;
        }
        // ------------------------- integrator update -------------------------
        float rx = ri.x, ry = ri.y, rz = ri.z;
        float vx = v4.x, vy = v4.y, vz = v4.z;
        const float m = v4.w;
        const float dt = P.dt;
        float K, fsq;
        {
#pragma clang fp contract(off)
            fsq = fx * fx;
            fsq = fsq + fy * fy;
            fsq = fsq + fz * fz;
            if (INTEG == 1) {  // NVE.py:78-97
                float vo, vm;
                vo = vx; vx = vx + (fx / m) * dt; vm = (vo + vx) / 2.0f; K = ((0.5f * m) * vm) * vm; rx = rx + vx * dt;
                vo = vy; vy = vy + (fy / m) * dt; vm = (vo + vy) / 2.0f; K = K + ((0.5f * m) * vm) * vm; ry = ry + vy * dt;
                vo = vz; vz = vz + (fz / m) * dt; vm = (vo + vz) / 2.0f; K = K + ((0.5f * m) * vm) * vm; rz = rz + vz * dt;
            } else if (INTEG == 2) {  // NVT.py:196-215
                const float factor = (0.5f * I.xi) * dt;
                const float plus = 1.0f / (1.0f + factor);
                const float minus = 1.0f - factor;
                vx = plus * (minus * vx + (fx / m) * dt); K = ((0.5f * m) * vx) * vx; rx = rx + vx * dt;
                vy = plus * (minus * vy + (fy / m) * dt); K = K + ((0.5f * m) * vy) * vy; ry = ry + vy * dt;
                vz = plus * (minus * vz + (fz / m) * dt); K = K + ((0.5f * m) * vz) * vz; rz = rz + vz * dt;
            } else {  // NVT_Langevin.py:363-378
                float z0, z1, z2;
                normal3(P.seed, P.ctrl->lstep0 + s, i, z0, z1, z2);
                float4 ob = P.obeta[i];
                const float num = 2.0f * m - P.alpha * dt;
                const float den = 2.0f * m + P.alpha * dt;
                const float a = num / den;
                const float bom = 2.0f / den;
                float b;
                b = I.lang_amp * z0; K = ((0.5f * m) * vx) * vx;
                vx = (a * vx + (bom * fx) * dt) + (bom * 0.5f) * (b + ob.x); ob.x = b; rx = rx + vx * dt;
                b = I.lang_amp * z1; K = K + ((0.5f * m) * vy) * vy;
                vy = (a * vy + (bom * fy) * dt) + (bom * 0.5f) * (b + ob.y); ob.y = b; ry = ry + vy * dt;
                b = I.lang_amp * z2; K = K + ((0.5f * m) * vz) * vz;
                vz = (a * vz + (bom * fz) * dt) + (bom * 0.5f) * (b + ob.z); ob.z = b; rz = rz + vz * dt;
                if (writer) P.obeta[i] = ob;
            }
        }
        int imx = 0, imy = 0, imz = 0;
        wrap_pbc(rx, imx, L0);
        wrap_pbc(ry, imy, L1);
        wrap_pbc(rz, imz, L2);
        if (writer) {
            posn[i] = make_float4(rx, ry, rz, ri.w);
            P.vel[i] = make_float4(vx, vy, vz, m);
            if (P.compute_k) P.sc[(size_t)SC_K * N + i] = K;
            if (P.compute_fsq) P.sc[(size_t)SC_FSQ * N + i] = fsq;
            if (imx | imy | imz) {
                P.rim[i] += imx;
                P.rim[(size_t)N + i] += imy;
                P.rim[(size_t)2 * N + i] += imz;
            }
            // skin test for the next interaction evaluation (nblist.py:71-73)
            const float4 rr = P.rref[i];
            float dx, dy, dz;
            const float d2 = dist_sq_mic(rx, ry, rz, rr.x, rr.y, rr.z, L0, L1, L2, dx, dy, dz);
            if (d2 > P.skin2q) P.rebuild_flag[s + 1] = 1;
            acc[RED_K] += K;
            acc[RED_FSQ] += fsq;
            {
#pragma clang fp contract(off)
                acc[RED_PX] += m * vx;
                acc[RED_PX + 1] += m * vy;
                acc[RED_PX + 2] += m * vz;
            }
            acc[RED_M] += m;
        }
    }
}

// Block-level reduction of the per-thread sums into partial[block] (fixed order).
__device__ inline void block_partials(const DP &P, float acc[GMD_NRED], double *red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int c = 0; c < GMD_NRED; ++c) {
        float v = wave_sum_f(acc[c]);
        if (lane == 0) red[wid * GMD_NRED + c] = (double)v;
    }
    __syncthreads();
    if (threadIdx.x < GMD_NRED) {
        double sum = 0.0;
        for (int w = 0; w < nw; ++w) sum += red[w * GMD_NRED + threadIdx.x];
        P.partial[(size_t)blockIdx.x * GMD_NRED + threadIdx.x] = (float)sum;
        __threadfence();
    }
    __syncthreads();
}

// Executed by the last block of a step: final sums over all blocks, then the
// one-thread updates of the reference (NVT.py:226-237 thermostat,
// momentum_reset.py:166-186 shift, scalar_saver.py:374-408 output row).
__device__ inline void step_tail(const DP &P, const StepInfo &I, double *red) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double acc[GMD_NRED];
#pragma unroll
    for (int c = 0; c < GMD_NRED; ++c) acc[c] = 0.0;
    if (I.need_ksum || I.need_mom || I.need_save) {
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) {
#pragma unroll
            for (int c = 0; c < GMD_NRED; ++c) acc[c] += (double)P.partial[(size_t)b * GMD_NRED + c];
        }
#pragma unroll
        for (int c = 0; c < GMD_NRED; ++c) {
            double v = wave_sum_d(acc[c]);
            if (lane == 0) red[wid * GMD_NRED + c] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot[GMD_NRED];
        for (int c = 0; c < GMD_NRED; ++c) {
            tot[c] = 0.0;
            if (I.need_ksum || I.need_mom || I.need_save)
                for (int w = 0; w < nw; ++w) tot[c] += red[w * GMD_NRED + c];
        }
        Ctrl *C = P.ctrl;
        if (I.need_ksum) {
            const float ksum = (float)tot[RED_K];
            const double t = (double)C->t0 + (double)I.s * (double)P.dt;
            const double T = target_temperature(P, t);
            const double dev = (double)(2.0f * ksum) / (P.degrees * T) - 1.0;
            const double inc = (double)(P.dt * P.omega2) * dev;
            C->xi = (float)((double)C->xi + inc);
            C->ksum = 0.0f;
        }
        if (I.need_mom) {
            const float mt = (float)tot[RED_M];
            C->shift[0] = (float)tot[RED_PX] / mt;
            C->shift[1] = (float)tot[RED_PX + 1] / mt;
            C->shift[2] = (float)tot[RED_PX + 2] / mt;
            C->shift_step = I.s + 1;
        }
        if (I.need_save) {
            float *row = P.saver_out + (size_t)(I.s / P.save_every) * 6;
            row[0] = (float)tot[RED_U];
            row[1] = (float)tot[RED_W];
            row[2] = (float)tot[RED_LAP];
            row[3] = (float)tot[RED_K];
            row[4] = (float)tot[RED_FSQ];
            row[5] = P.box0 * P.box1 * P.box2;
        }
        __threadfence();
    }
    __syncthreads();
}

// Software grid barrier for the cooperative (co-resident) launch.  The last
// block to arrive may run a tail function before releasing the others.
template <typename Tail>
__device__ inline void grid_barrier(Ctrl *C, Tail tail) {
    __shared__ int s_last;
    __shared__ unsigned s_gen;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned gen = __hip_atomic_load(&C->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        unsigned arrived = __hip_atomic_fetch_add(&C->bar_count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = (arrived == gridDim.x - 1);
        s_gen = gen;
    }
    __syncthreads();
    if (s_last) {
        __threadfence();
        tail();
        if (threadIdx.x == 0) {
            __hip_atomic_store(&C->bar_count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&C->bar_gen, s_gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    } else if (threadIdx.x == 0) {
        while (__hip_atomic_load(&C->bar_gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == s_gen)
            __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
    if (threadIdx.x == 0) __threadfence();
    __syncthreads();
}

struct NoTail {
    __device__ void operator()() const {}
};
