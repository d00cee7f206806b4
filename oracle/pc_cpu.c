/*
 * CPU restatement of the layer algorithms on the compiled index tables.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py load it.
 *
 * Built twice by oracle/build_cpu.py: with PCC_REAL=double (an FP64 oracle on the compiled
 * circuit, fast enough for the full-size circuits) and PCC_REAL=float (the FP32 CPU baseline),
 * both with OpenMP over node blocks.  Index tables are the compiler's int64 partition tensors;
 * buffers are row-major [rows][B] like the device buffers.
 *
 * Each function follows the reference layer it is named after (paths in the PyJuice tree):
 *   pcc_input_forward   layer/input_layer.py:283-404, categorical.py:83-92, bernoulli.py:65-71
 *   pcc_input_backward  layer/input_layer.py:406-571, categorical.py:94-101, bernoulli.py:73-77
 *   pcc_input_em        layer/input_layer.py:663-724, categorical.py:164-245, bernoulli.py:91-101
 *   pcc_prod            layer/prod_layer.py:398-466 (forward and flow backward gather-add)
 *   pcc_sum_forward     layer/sum_layer.py:454-568
 *   pcc_sum_ele         layer/sum_layer.py:1569-1730 (element flows, unmodified node flows)
 *   pcc_sum_par         layer/sum_layer.py:2049-2169 (parameter flows)
 *   pcc_sum_em          model/backend/par_update.py:173-256
 *   pcc_pflow_fuse      model/backend/parflow_fusing.py:81-112
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#ifndef PCC_REAL
#define PCC_REAL double
#endif
typedef PCC_REAL real;

#define NEG_INF (-INFINITY)

/* node_mars[off + i][b] = log p_i(x[vids[i]][b]) */
void pcc_input_forward(int kind, const real* params, real* node_mars, const int32_t* data, const int64_t* vids,
                       const int64_t* s_pids, const int64_t* ncats, int64_t num_nodes, int64_t off, int64_t B) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < num_nodes; ++i) {
        const int32_t* x = data + vids[i] * B;
        real* out = node_mars + (off + i) * B;
        for (int64_t b = 0; b < B; ++b) {
            if (kind == 0) {
                const int64_t k = x[b];
                out[b] = (k >= 0 && k < ncats[i]) ? log(params[s_pids[i] + k]) : NEG_INF;
            } else {
                const real p = params[s_pids[i]];
                out[b] = log(x[b] > 0 ? p : (real)1 - p);
            }
        }
    }
}

/* Categorical: pf[s_pfids[i] + x] += flow; Bernoulli: slot 0 collects x = 1, slot 1 x = 0. */
void pcc_input_backward(int kind, real* pflows, const real* node_flows, const int32_t* data, const int64_t* vids,
                        const int64_t* s_pfids, const int64_t* ncats, int64_t num_nodes, int64_t off, int64_t B) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < num_nodes; ++i) {
        const int32_t* x = data + vids[i] * B;
        const real* f = node_flows + (off + i) * B;
        for (int64_t b = 0; b < B; ++b) {
            int64_t slot;
            if (kind == 0) {
                if (x[b] < 0 || x[b] >= ncats[i]) continue;
                slot = s_pfids[i] + x[b];
            } else {
                slot = s_pfids[i] + (x[b] > 0 ? 0 : 1);
            }
#pragma omp atomic
            pflows[slot] += f[b];
        }
    }
}

/* theta <- (1 - s) theta + s (f + pc / K) / (sum f + pc) for the source nodes src[].  keep_zero
 * (Categorical): parameters below 1e-12 stay 0 and the normaliser is sum over the others of
 * (f + pc / K) (categorical.py:171-201). */
void pcc_input_em(int kind, real* params, const real* pflows, const int64_t* src, const int64_t* s_pids,
                  const int64_t* s_pfids, const int64_t* ncats, int64_t num_src, double step, double pc,
                  int keep_zero) {
#pragma omp parallel for schedule(static)
    for (int64_t w = 0; w < num_src; ++w) {
        const int64_t i = src[w];
        if (kind == 0) {
            const int64_t K = ncats[i];
            const double npc = pc / (double)K;
            double cum = keep_zero ? 0.0 : pc;
            for (int64_t k = 0; k < K; ++k) {
                const double f = pflows[s_pfids[i] + k];
                if (!keep_zero) cum += f;
                else if (params[s_pids[i] + k] >= 1e-12) cum += f + npc;
            }
            for (int64_t k = 0; k < K; ++k) {
                const double th = params[s_pids[i] + k];
                const double nv = ((double)pflows[s_pfids[i] + k] + npc) / cum;
                params[s_pids[i] + k] = (keep_zero && th < 1e-12) ? (real)0 : (real)((1.0 - step) * th + step * nv);
            }
        } else {
            const double f1 = pflows[s_pfids[i]], f0 = pflows[s_pfids[i] + 1];
            const double nv = (f1 + pc / 2.0) / (f1 + f0 + pc);
            params[s_pids[i]] = (real)((1.0 - step) * params[s_pids[i]] + step * nv);
        }
    }
}

/* out[nids[j] + r][b] (+)= sum_k in[cids[j][k] + r][b], cids == 0 is padding. */
void pcc_prod(real* out, const real* in, const int64_t* nids, const int64_t* cids, int64_t n_blocks, int64_t K,
              int64_t bs, int64_t B, int accum) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n_blocks * bs; ++t) {
        const int64_t j = t / bs, r = t % bs;
        real* o = out + (nids[j] + r) * B;
        if (!accum)
            for (int64_t b = 0; b < B; ++b) o[b] = 0;
        for (int64_t k = 0; k < K; ++k) {
            const int64_t c = cids[j * K + k];
            if (c == 0) continue;
            const real* src = in + (c + r) * B;
            for (int64_t b = 0; b < B; ++b) o[b] += src[b];
        }
    }
}

/* The three sum-layer passes below factor every exp(x - y) through a shift per (node block,
 * column): the exponentials are taken once per (child row, column) and once per (parent row,
 * column), and the contraction over the block's rows is a plain multiply-add, so the cost is that
 * of a dense product rather than one exponential per (edge, sample).  Columns are processed in
 * chunks of PCC_BC; OpenMP runs over (block, chunk) tasks.  Where a shifted exponent would
 * overflow (possible only next to zero parameters), the term is summed directly instead. */
#define PCC_BC 64
#define PCC_EXP_MAX ((sizeof(real) == sizeof(double)) ? 700.0 : 80.0)

static inline real pcc_exp_or_zero(real x) { return (x == NEG_INF) ? (real)0 : (real)exp(x); }

/* node_mars[nids[j] + r][b] = log sum_k theta[pids[j][k] + r] exp(e[cids[j][k]][b]) over the
 * children k < nch[j]; per column the children's maximum is subtracted before the exponential
 * (which is then independent of r). */
void pcc_sum_forward(real* node_mars, const real* emars, const real* params, const int64_t* nids,
                     const int64_t* cids, const int64_t* pids, const int64_t* nch, int64_t n_blocks, int64_t K,
                     int64_t bs, int64_t B) {
    const int64_t nc = (B + PCC_BC - 1) / PCC_BC;
#pragma omp parallel
    {
        real* acc = (real*)malloc(sizeof(real) * bs * PCC_BC);
        real mx[PCC_BC], ex[PCC_BC];
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < n_blocks * nc; ++t) {
            const int64_t j = t / nc, b0 = (t % nc) * PCC_BC;
            const int64_t nb = (B - b0) < PCC_BC ? (B - b0) : PCC_BC;
            const int64_t n = nch[j];
            for (int64_t b = 0; b < nb; ++b) mx[b] = NEG_INF;
            for (int64_t k = 0; k < n; ++k) {
                const real* e = emars + cids[j * K + k] * B + b0;
                for (int64_t b = 0; b < nb; ++b) if (e[b] > mx[b]) mx[b] = e[b];
            }
            for (int64_t q = 0; q < bs * nb; ++q) acc[q] = 0;
            for (int64_t k = 0; k < n; ++k) {
                const real* e = emars + cids[j * K + k] * B + b0;
                for (int64_t b = 0; b < nb; ++b) ex[b] = (mx[b] == NEG_INF) ? (real)0 : pcc_exp_or_zero(e[b] - mx[b]);
                const real* th = params + pids[j * K + k];
                for (int64_t r = 0; r < bs; ++r) {
                    const real w = th[r];
                    real* a = acc + r * nb;
                    for (int64_t b = 0; b < nb; ++b) a[b] += w * ex[b];
                }
            }
            for (int64_t r = 0; r < bs; ++r) {
                real* o = node_mars + (nids[j] + r) * B + b0;
                const real* a = acc + r * nb;
                for (int64_t b = 0; b < nb; ++b)
                    o[b] = (a[b] > 0 && mx[b] != NEG_INF) ? (real)log(a[b]) + mx[b] : NEG_INF;
            }
        }
        free(acc);
    }
}

/* R[i][b] = sign(f) exp(log|f| - m - s[b]) for the rows i < bs starting at row0, with
 * s[b] = max_i (log|f| - m) (-inf where no row has flow). */
static void pcc_scaled_flows(real* R, real* s, const real* nflows, const real* nmars, int64_t row0, int64_t bs,
                             int64_t b0, int64_t nb, int64_t B) {
    for (int64_t b = 0; b < nb; ++b) s[b] = NEG_INF;
    for (int64_t i = 0; i < bs; ++i) {
        const real* f = nflows + (row0 + i) * B + b0;
        const real* m = nmars + (row0 + i) * B + b0;
        real* l = R + i * nb;
        for (int64_t b = 0; b < nb; ++b) {
            l[b] = (f[b] != 0 && m[b] != NEG_INF) ? (real)log(fabs(f[b])) - m[b] : NEG_INF;
            if (l[b] > s[b]) s[b] = l[b];
        }
    }
    for (int64_t i = 0; i < bs; ++i) {
        const real* f = nflows + (row0 + i) * B + b0;
        real* l = R + i * nb;
        for (int64_t b = 0; b < nb; ++b) {
            const real v = (l[b] == NEG_INF) ? (real)0 : (real)exp(l[b] - s[b]);
            l[b] = f[b] < 0 ? -v : v;
        }
    }
}

/* eflows[chids[j] + c][b] (+)= sum_{k < npar[j], i < bs} theta[parpids[j][k] + c bs + i]
 *                              * f[parids[j][k] + i][b] * exp(e[chids[j] + c][b] - m[parids[j][k] + i][b])
 * (node flows unmodified, any sign). */
void pcc_sum_ele(real* eflows, const real* nflows, const real* nmars, const real* emars, const real* params,
                 const int64_t* chids, const int64_t* parids, const int64_t* parpids, const int64_t* npar, int64_t n,
                 int64_t KP, int64_t cbs, int64_t bs, int64_t B, int accum) {
    const int64_t nc = (B + PCC_BC - 1) / PCC_BC;
    const real lim = (real)PCC_EXP_MAX;
#pragma omp parallel
    {
        real* R = (real*)malloc(sizeof(real) * bs * PCC_BC);
        real s[PCC_BC], acc[PCC_BC];
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < n * nc; ++t) {
            const int64_t j = t / nc, b0 = (t % nc) * PCC_BC;
            const int64_t nb = (B - b0) < PCC_BC ? (B - b0) : PCC_BC;
            if (!accum)
                for (int64_t c = 0; c < cbs; ++c)
                    for (int64_t b = 0; b < nb; ++b) eflows[(chids[j] + c) * B + b0 + b] = 0;
            for (int64_t k = 0; k < npar[j]; ++k) {
                const int64_t prow = parids[j * KP + k];
                pcc_scaled_flows(R, s, nflows, nmars, prow, bs, b0, nb, B);
                for (int64_t c = 0; c < cbs; ++c) {
                    const int64_t row = chids[j] + c;
                    const real* e = emars + row * B + b0;
                    const real* th = params + parpids[j * KP + k] + c * bs;
                    for (int64_t b = 0; b < nb; ++b) acc[b] = 0;
                    for (int64_t i = 0; i < bs; ++i) {
                        const real w = th[i];
                        const real* r = R + i * nb;
                        for (int64_t b = 0; b < nb; ++b) acc[b] += w * r[b];
                    }
                    real* o = eflows + row * B + b0;
                    for (int64_t b = 0; b < nb; ++b) {
                        if (e[b] == NEG_INF || s[b] == NEG_INF) continue;
                        if (e[b] + s[b] <= lim) {
                            o[b] += acc[b] * (real)exp(e[b] + s[b]);
                        } else {   /* direct sum: the shifted exponent would overflow */
                            real d = 0;
                            for (int64_t i = 0; i < bs; ++i) {
                                const real f = nflows[(prow + i) * B + b0 + b];
                                const real m = nmars[(prow + i) * B + b0 + b];
                                if (f != 0 && m != NEG_INF) d += th[i] * f * (real)exp(e[b] - m);
                            }
                            o[b] += d;
                        }
                    }
                }
            }
        }
        free(R);
    }
}

/* pflows[pfids[j][k] + r] += theta[pids[j][k] + r] * sum_b f[nids[j] + r][b]
 *                            * exp(e[cids[j][k]][b] - m[nids[j] + r][b])  (atomic: slots may be shared) */
void pcc_sum_par(real* pflows, const real* nflows, const real* nmars, const real* emars, const real* params,
                 const int64_t* nids, const int64_t* cids, const int64_t* pids, const int64_t* pfids,
                 const int64_t* nch, int64_t n_blocks, int64_t K, int64_t bs, int64_t B) {
    const int64_t nc = (B + PCC_BC - 1) / PCC_BC;
    const real lim = (real)PCC_EXP_MAX;
#pragma omp parallel
    {
        real* R = (real*)malloc(sizeof(real) * bs * PCC_BC);
        real* loc = (real*)malloc(sizeof(real) * bs);
        real s[PCC_BC], C[PCC_BC];
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < n_blocks * nc; ++t) {
            const int64_t j = t / nc, b0 = (t % nc) * PCC_BC;
            const int64_t nb = (B - b0) < PCC_BC ? (B - b0) : PCC_BC;
            pcc_scaled_flows(R, s, nflows, nmars, nids[j], bs, b0, nb, B);
            for (int64_t k = 0; k < nch[j]; ++k) {
                const real* e = emars + cids[j * K + k] * B + b0;
                int direct = 0;
                for (int64_t b = 0; b < nb; ++b) {
                    if (e[b] == NEG_INF || s[b] == NEG_INF) { C[b] = 0; continue; }
                    if (e[b] + s[b] > lim) { C[b] = 0; direct = 1; continue; }
                    C[b] = (real)exp(e[b] + s[b]);
                }
                for (int64_t r = 0; r < bs; ++r) {
                    const real* rr = R + r * nb;
                    real a = 0;
                    for (int64_t b = 0; b < nb; ++b) a += rr[b] * C[b];
                    loc[r] = a;
                }
                if (direct) {   /* columns whose shifted exponent would overflow: direct sum */
                    for (int64_t b = 0; b < nb; ++b) {
                        if (e[b] == NEG_INF || s[b] == NEG_INF || e[b] + s[b] <= lim) continue;
                        for (int64_t r = 0; r < bs; ++r) {
                            const real f = nflows[(nids[j] + r) * B + b0 + b];
                            const real m = nmars[(nids[j] + r) * B + b0 + b];
                            if (f != 0 && m != NEG_INF) loc[r] += f * (real)exp(e[b] - m);
                        }
                    }
                }
                const real* th = params + pids[j * K + k];
                real* pf = pflows + pfids[j * K + k];
                for (int64_t r = 0; r < bs; ++r) {
                    const real v = th[r] * loc[r];
#pragma omp atomic
                    pf[r] += v;
                }
            }
        }
        free(loc);
        free(R);
    }
}

/* Sum EM over node blocks rec = {par_start, pflow_start, node_start, bs, nch}: the parameter of
 * child c of node i is at par_start + c bs + i.  keep_zero: parameters below 1e-12 stay 0 and the
 * normaliser is the sum over the others of (pf + pc / nch) (par_update.py:196-205, 236-253). */
void pcc_sum_em(real* params, const real* pflows, const int64_t* recs, int64_t n_recs, double step, double pc,
                int keep_zero) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t q = 0; q < n_recs; ++q) {
        const int64_t* R = recs + q * 5;
        const int64_t par0 = R[0], pf0 = R[1], bs = R[3], nch = R[4];
        const double npc = pc / (double)nch;
        for (int64_t i = 0; i < bs; ++i) {
            double cum = keep_zero ? 0.0 : pc;
            for (int64_t c = 0; c < nch; ++c) {
                const double f = pflows[pf0 + c * bs + i];
                if (!keep_zero) cum += f;
                else if (params[par0 + c * bs + i] >= 1e-12) cum += f + npc;
            }
            for (int64_t c = 0; c < nch; ++c) {
                const int64_t o = c * bs + i;
                const double th = params[par0 + o];
                const double nv = ((double)pflows[pf0 + o] + npc) / cum;
                params[par0 + o] = (keep_zero && th < 1e-12) ? (real)0 : (real)((1.0 - step) * th + step * nv);
            }
        }
    }
}

/* pf[target + j] = sum_c pf[copies[c] + j] (the copies include the target). */
void pcc_pflow_fuse(real* pflows, int64_t target, int64_t n, const int64_t* copies, int64_t ncopies) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        real s = 0;
        for (int64_t c = 0; c < ncopies; ++c) s += pflows[copies[c] + j];
        pflows[target + j] = s;
    }
}
