/*
 * cpu_md.c -- C/OpenMP restatement of gamdpy's per-timestep path.
 * TEST INFRASTRUCTURE ONLY: used by the tests as a checker for sizes NumPy
 * cannot reach in seconds, and timed by bench.py as the CPU baseline.
 * Never linked into or called by the product library.
 *
 * Follows (reference files, relative to the gamdpy checkout):
 *   simulation_boxes/orthorhombic.py:43-132      minimum image, PBC, skin test
 *   interactions/nblist_linked_lists.py:36-193   cell lists, 5x5x5 stencil
 *   interactions/pair_potential.py:162-246       full-list force calculator
 *   potential_functions/LJ_12_6_sigma_epsilon.py:25-36, apply_shifted_potential_cutoff.py:41-47
 *   integrators/NVT.py:188-237, integrators/NVE.py:59-105
 *   calculators/calculator_radial_distribution.py:76-109
 * Built with -ffp-contract=off so every float operation is rounded once, in
 * the order written (the same order as oracle/md_oracle.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float mic1(float d, float L) {
    float t = 2.0f * d;
    return t > L ? d - L : (t < -L ? d + L : d);
}

static inline float dist_sq_dr(const float *a, const float *b, const float *L, float *dr) {
    dr[0] = mic1(a[0] - b[0], L[0]);
    dr[1] = mic1(a[1] - b[1], L[1]);
    dr[2] = mic1(a[2] - b[2], L[2]);
    float d2 = dr[0] * dr[0];
    d2 = d2 + dr[1] * dr[1];
    d2 = d2 + dr[2] * dr[2];
    return d2;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Radial distribution counts: ordered pairs, f32 distance, f64 binning. */
int orc_rdf_counts(int N, const float *r, const int *ptype, const float *box, int nt, int bins,
                   long long *counts) {
    float mn = box[0] < box[1] ? box[0] : box[1];
    mn = mn < box[2] ? mn : box[2];
    const double half = (double)mn / 2.0;
    const double bw = half / bins;
    const size_t nh = (size_t)nt * nt * bins;
    memset(counts, 0, nh * sizeof(long long));
#pragma omp parallel
    {
        long long *loc = (long long *)calloc(nh, sizeof(long long));
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < N; ++i) {
            float dr[3];
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                float d2 = dist_sq_dr(r + 3 * j, r + 3 * i, box, dr);
                if (!((double)d2 < half * half)) continue;
                float d = sqrtf(d2);
                if (!((double)d < half)) continue;
                int b = (int)floor((double)d / bw);
                loc[((size_t)ptype[i] * nt + ptype[j]) * bins + b] += 1;
            }
        }
#pragma omp critical
        for (size_t x = 0; x < nh; ++x) counts[x] += loc[x];
        free(loc);
    }
    return 0;
}

/* --------------------------------------------------------------------------
 * Cell-list neighbour list (full list), nblist_linked_lists.py:36-193.
 * list: N x max_nbs (row-major), cnt: N (true counts).  Returns max count.
 * ------------------------------------------------------------------------ */
int orc_build_ll(int N, const float *r, const float *box, float max_cut, float skin, int max_nbs, int *list,
                 int *cnt) {
    const double min_size = ((double)max_cut + (double)skin) / 2.0;
    int nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = (int)((double)box[k] / min_size);
    if (nc[0] <= 6 || nc[1] <= 4 || nc[2] <= 4) return -1;
    const int ncells = nc[0] * nc[1] * nc[2];
    int *head = (int *)malloc(sizeof(int) * ncells);
    int *next = (int *)malloc(sizeof(int) * N);
    int *cell = (int *)malloc(sizeof(int) * 3 * N);
    for (int c = 0; c < ncells; ++c) head[c] = -1;
    for (int i = 0; i < N; ++i) {
        for (int k = 0; k < 3; ++k) {
            long long ck = (long long)floor(((double)r[3 * i + k] + 0.5 * (double)box[k]) * nc[k] / (double)box[k]);
            ck %= nc[k];
            if (ck < 0) ck += nc[k];
            cell[3 * i + k] = (int)ck;
        }
        int c = (cell[3 * i] * nc[1] + cell[3 * i + 1]) * nc[2] + cell[3 * i + 2];
        next[i] = head[c];
        head[c] = i;
    }
    float cps = max_cut + skin;
    const float thr = cps * cps;
    int maxc = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : maxc)
    for (int i = 0; i < N; ++i) {
        int n = 0;
        float dr[3];
        for (int ix = -2; ix <= 2; ++ix)
            for (int iy = -2; iy <= 2; ++iy)
                for (int iz = -2; iz <= 2; ++iz) {
                    int ox = ((cell[3 * i] + ix) % nc[0] + nc[0]) % nc[0];
                    int oy = ((cell[3 * i + 1] + iy) % nc[1] + nc[1]) % nc[1];
                    int oz = ((cell[3 * i + 2] + iz) % nc[2] + nc[2]) % nc[2];
                    for (int j = head[(ox * nc[1] + oy) * nc[2] + oz]; j >= 0; j = next[j]) {
                        if (j == i) continue;
                        if (dist_sq_dr(r + 3 * j, r + 3 * i, box, dr) < thr) {
                            if (n < max_nbs) list[(size_t)i * max_nbs + n] = j;
                            ++n;
                        }
                    }
                }
        cnt[i] = n;
        if (n > maxc) maxc = n;
    }
    free(head);
    free(next);
    free(cell);
    return maxc;
}

/* LJ sigma/epsilon with the shifted-potential cutoff, full-list calculator. */
static inline void lj_se(float dist, float sigma, float eps, float *u, float *s, float *umm) {
    float od = 1.0f / dist;
    float x = sigma * od;
    float x2 = x * x;
    float x6 = x2 * x2 * x2;
    float x12 = x6 * x6;
    float od2 = od * od;
    *u = (4.0f * eps) * (x12 - x6);
    *s = ((24.0f * eps) * (2.0f * x12 - x6)) * od2;
    *umm = ((24.0f * eps) * (26.0f * x12 - 7.0f * x6)) * od2;
}

int orc_forces_sp(int N, const float *r, const int *ptype, const float *box, int nt, const float *params,
                  const int *list, const int *cnt, int max_nbs, float *f, float *U, float *W) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        float fx = 0.f, fy = 0.f, fz = 0.f, u_i = 0.f, w_i = 0.f;
        float dr[3];
        const int n = cnt[i] < max_nbs ? cnt[i] : max_nbs;
        for (int k = 0; k < n; ++k) {
            const int j = list[(size_t)i * max_nbs + k];
            const float *pp = params + 3 * (ptype[i] * nt + ptype[j]);
            float d2 = dist_sq_dr(r + 3 * j, r + 3 * i, box, dr);
            float cut = pp[2];
            if (d2 < cut * cut) {
                float u, s, umm, uc, sc, ummc;
                float dist = sqrtf(d2);
                lj_se(dist, pp[0], pp[1], &u, &s, &umm);
                lj_se(cut, pp[0], pp[1], &uc, &sc, &ummc);
                u = u - uc;
                fx = fx - dr[0] * s;
                fy = fy - dr[1] * s;
                fz = fz - dr[2] * s;
                w_i = w_i + ((dr[0] * dr[0]) * s) * (0.5f / 3.0f);
                w_i = w_i + ((dr[1] * dr[1]) * s) * (0.5f / 3.0f);
                w_i = w_i + ((dr[2] * dr[2]) * s) * (0.5f / 3.0f);
                u_i = u_i + 0.5f * u;
            }
        }
        f[3 * i] = fx;
        f[3 * i + 1] = fy;
        f[3 * i + 2] = fz;
        U[i] = u_i;
        W[i] = w_i;
    }
    return 0;
}

/*
 * NVT (Nose-Hoover leapfrog) with linked lists: 'steps' steps, steps+1
 * interaction evaluations (Simulation.py:312-324).  xi_state = [xi, sum K].
 * Returns the number of list rebuilds, or -1 on error (too few cells,
 * neighbour overflow).
 */
long long orc_md_nvt_ll(int N, float *r, float *v, const float *m, const int *ptype, int *img, const float *box,
                        int nt, const float *params, float max_cut, float skin, int max_nbs, float dt, float T,
                        float omega2, long long degrees, float *xi_state, int steps, float *f, float *U, float *W,
                        float *K) {
    int *list = (int *)malloc(sizeof(int) * (size_t)N * max_nbs);
    int *cnt = (int *)malloc(sizeof(int) * N);
    float *rref = (float *)calloc((size_t)3 * N, sizeof(float));
    long long rebuilds = 0;
    const float skin2q = (skin * skin) * 0.25f;
    float xi = xi_state[0];
    int need = 1;
    for (int step = 0; step <= steps; ++step) {
        if (!need) {
            int any = 0;
#pragma omp parallel for reduction(| : any)
            for (int i = 0; i < N; ++i) {
                float dr[3];
                if (dist_sq_dr(r + 3 * i, rref + 3 * i, box, dr) > skin2q) any |= 1;
            }
            need = any;
        }
        if (need) {
            int mx = orc_build_ll(N, r, box, max_cut, skin, max_nbs, list, cnt);
            if (mx < 0 || mx > max_nbs) {
                free(list); free(cnt); free(rref);
                return -1;
            }
            memcpy(rref, r, sizeof(float) * 3 * (size_t)N);
            ++rebuilds;
            need = 0;
        }
        orc_forces_sp(N, r, ptype, box, nt, params, list, cnt, max_nbs, f, U, W);
        if (step == steps) break;
        const float factor = (0.5f * xi) * dt;
        const float plus = 1.0f / (1.0f + factor);
        const float minus = 1.0f - factor;
        double ksum = 0.0;
#pragma omp parallel for reduction(+ : ksum)
        for (int i = 0; i < N; ++i) {
            float k_i = 0.f;
            for (int k = 0; k < 3; ++k) {
                float vk = plus * (minus * v[3 * i + k] + (f[3 * i + k] / m[i]) * dt);
                v[3 * i + k] = vk;
                k_i = k_i + ((0.5f * m[i]) * vk) * vk;
                float rk = r[3 * i + k] + vk * dt;
                if (rk * 2.0f > box[k]) { rk = rk - box[k]; img[3 * i + k] += 1; }
                if (rk * 2.0f < -box[k]) { rk = rk + box[k]; img[3 * i + k] -= 1; }
                r[3 * i + k] = rk;
            }
            K[i] = k_i;
            ksum += k_i;
        }
        const double dev = (double)(2.0f * (float)ksum) / ((double)degrees * (double)T) - 1.0;
        xi = (float)((double)xi + (double)(dt * omega2) * dev);
    }
    xi_state[0] = xi;
    xi_state[1] = 0.0f;
    free(list);
    free(cnt);
    free(rref);
    return rebuilds;
}
