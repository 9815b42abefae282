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

/*
 * Differentiable-trajectory observable in float64 (the checker of the
 * adjoint path; finite differences of it pin the adjoint's gradients).
 * LJ_12_6_sigma_epsilon with the shifted-force cutoff at a fixed cut
 * (apply_shifted_force_cutoff.py:29-37, s -= s(rc)): s(r) = 24 eps (2 (sig/r)^12 - (sig/r)^6) / r^2 - s(rc);
 * leap-frog NVE (NVE.py:59-105) with
 * the PBC wrap of orthorhombic.py; histogram H[c][k] summed over frames
 * n % rdf_every == 0 (n = 0..nsteps) and unordered pairs of class c of
 * max(0, exp(-(d - r_k)^2 / 2w^2) - exp(-12.5)) for |d - r_k| < 5w, r_k = (k + 1/2) rmax / bins.
 * params: nt*nt*3 (sigma, epsilon, cut); classes a <= b in row order.
 */
static inline double mic1d(double d, double L) {
    double t = 2.0 * d;
    return t > L ? d - L : (t < -L ? d + L : d);
}

static int adj_cls(int a, int b, int nt) {
    if (a > b) {
        int t = a;
        a = b;
        b = t;
    }
    return a * nt - a * (a - 1) / 2 + (b - a);
}

static void adj_hist_frame(int N, const double *r, const int *ptype, const double *box, int nt, int bins,
                           double rmax, double w, double *H) {
    const double dr = rmax / bins, rc = rmax + 5.0 * w, floor_ = exp(-12.5);
    const int ncls = nt * (nt + 1) / 2;
#pragma omp parallel
    {
        double *loc = (double *)calloc((size_t)ncls * bins, sizeof(double));
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < N; ++i)
            for (int j = i + 1; j < N; ++j) {
                double dx = mic1d(r[3 * i] - r[3 * j], box[0]);
                double dy = mic1d(r[3 * i + 1] - r[3 * j + 1], box[1]);
                double dz = mic1d(r[3 * i + 2] - r[3 * j + 2], box[2]);
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 >= rc * rc) continue;
                double d = sqrt(d2);
                int c = adj_cls(ptype[i], ptype[j], nt);
                for (int k = 0; k < bins; ++k) {
                    double x = d - (k + 0.5) * dr;
                    if (fabs(x) >= 5.0 * w) continue;
                    double kv = exp(-0.5 * x * x / (w * w)) - floor_;
                    if (kv > 0.0) loc[(size_t)c * bins + k] += kv;
                }
            }
#pragma omp critical
        for (size_t k = 0; k < (size_t)ncls * bins; ++k) H[k] += loc[k];
        free(loc);
    }
}

int orc_adj_histogram_f64(int N, const double *r0, const double *v0, const double *mass, const int *ptype,
                          const double *box, int nt, const double *params, double dt, int nsteps, int rdf_every,
                          int bins, double rmax, double width, double *H, double *r_end) {
    const int ncls = nt * (nt + 1) / 2;
    double *r = (double *)malloc(sizeof(double) * 3 * N);
    double *v = (double *)malloc(sizeof(double) * 3 * N);
    double *f = (double *)malloc(sizeof(double) * 3 * N);
    double sc[64];
    for (int a = 0; a < nt * nt; ++a) {
        double sig = params[3 * a], eps = params[3 * a + 1], rc = params[3 * a + 2];
        double x6 = pow(sig / rc, 6.0);
        sc[a] = 24.0 * eps * (2.0 * x6 * x6 - x6) / (rc * rc);
    }
    memcpy(r, r0, sizeof(double) * 3 * N);
    memcpy(v, v0, sizeof(double) * 3 * N);
    memset(H, 0, sizeof(double) * ncls * bins);
    for (int n = 0; n <= nsteps; ++n) {
        if (n % rdf_every == 0) adj_hist_frame(N, r, ptype, box, nt, bins, rmax, width, H);
        if (n == nsteps) break;
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < N; ++i) {
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int j = 0; j < N; ++j) {
                if (j == i) continue;
                double dx = mic1d(r[3 * i] - r[3 * j], box[0]);
                double dy = mic1d(r[3 * i + 1] - r[3 * j + 1], box[1]);
                double dz = mic1d(r[3 * i + 2] - r[3 * j + 2], box[2]);
                double d2 = dx * dx + dy * dy + dz * dz;
                const double *p = params + 3 * (ptype[i] * nt + ptype[j]);
                if (d2 >= p[2] * p[2]) continue;
                double x = p[0] * p[0] / d2, x3 = x * x * x, x6 = x3 * x3;
                double s = 24.0 * p[1] * (2.0 * x6 - x3) / d2 - sc[ptype[i] * nt + ptype[j]];
                fx += s * dx;
                fy += s * dy;
                fz += s * dz;
            }
            f[3 * i] = fx;
            f[3 * i + 1] = fy;
            f[3 * i + 2] = fz;
        }
        for (int i = 0; i < N; ++i)
            for (int k = 0; k < 3; ++k) {
                v[3 * i + k] += f[3 * i + k] / mass[i] * dt;
                double x = r[3 * i + k] + v[3 * i + k] * dt;
                if (2.0 * x > box[k]) x -= box[k];
                if (2.0 * x < -box[k]) x += box[k];
                r[3 * i + k] = x;
            }
    }
    if (r_end) memcpy(r_end, r, sizeof(double) * 3 * N);
    free(r);
    free(v);
    free(f);
    return 0;
}
