/*
 * CPU oracle for the field-evaluation path: a plain-C restatement of the reference's
 * Pyccel kernels (FP64, OpenMP where the reference has an `omp parallel for`).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this library, and only as the checker or the timed CPU
 * baseline.  The product path (tree-d-3d_amd/) never calls it.
 *
 * Every function cites the reference file it follows (paths relative to the reference's
 * src/tree_d_3d/kernels/).  Arrays follow the reference's layouts: (2, N) component-major
 * point arrays with an explicit row stride `ld` (the reference's tree-order copies carry
 * one spare column), node-major (C, k) per-node arrays, int64 indices.
 *
 * Pinning (what checks each function independently of its own formulas) is listed in
 * oracle/README.md and exercised by tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

#define TD3_BAD 1.0e300

/* ------------------------------------------------------------------------------------ */
/* direct_sum/direct_sum_pyccel.py:15-72 and direct_log_sum/direct_log_sum_pyccel.py:15-70 */
/* ------------------------------------------------------------------------------------ */
void oracle_direct(int log_kernel, const double *targets, i64 ldt, i64 m, const i64 *exclude,
                   const double *pos, i64 ldp, i64 n, const double *charge, double softening,
                   double *phi, double *field, i64 ldf, i64 *status) {
    const double eps2 = softening * softening;
    i64 num_bad = 0;
#pragma omp parallel for schedule(dynamic, 8) reduction(+ : num_bad)
    for (i64 i = 0; i < m; ++i) {
        const double tx = targets[i], ty = targets[ldt + i];
        const i64 skip = exclude[i];
        double potential = 0.0, ex = 0.0, ey = 0.0;
        int bad = 0;
        for (i64 j = 0; j < n; ++j) {
            const double c = charge[j];
            if (j == skip || c == 0.0) continue;
            const double dx = tx - pos[j];
            const double dy = ty - pos[ldp + j];
            const double r2 = dx * dx + dy * dy + eps2;
            if (r2 == 0.0) {
                bad = 1;
                break;
            }
            if (log_kernel) {
                potential -= c * log(r2);
                ex += 2.0 * c * dx / r2;
                ey += 2.0 * c * dy / r2;
            } else {
                const double inv_r = 1.0 / sqrt(r2);
                const double inv_r3 = inv_r * inv_r * inv_r;
                potential += c * inv_r;
                ex += c * inv_r3 * dx;
                ey += c * inv_r3 * dy;
            }
        }
        if (bad) {
            num_bad += 1;
            phi[i] = TD3_BAD;
        } else {
            phi[i] = potential;
            field[i] = ex;
            field[ldf + i] = ey;
        }
    }
    if (num_bad > 0) {
        for (i64 i = 0; i < m; ++i) {
            if (phi[i] == TD3_BAD) {
                phi[i] = 0.0;
                status[0] = i + 1;
                return;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* quadtree_build/quadtree_build_pyccel.py:12-137 (serial, breadth first, stable).       */
/* sorted_pos and new_sorted_pos are (2, ld) with ld >= n.                               */
/* ------------------------------------------------------------------------------------ */
void oracle_quadtree_build(double *sorted_pos, i64 ld, i64 n, i64 leaf_capacity, i64 max_depth,
                           double *node_center, double *node_half, i64 *node_start,
                           i64 *node_end, i64 *node_child, i64 *node_depth, i64 *perm,
                           i64 *levels, i64 capacity, i64 *sizes) {
    i64 *slot_node = (i64 *)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
    i64 *quadrant = (i64 *)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
    i64 *cursor = (i64 *)malloc(sizeof(i64) * (size_t)capacity);
    i64 *new_perm = (i64 *)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
    i64 *new_slot_node = (i64 *)malloc(sizeof(i64) * (size_t)(n > 0 ? n : 1));
    double *new_pos = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    for (i64 k = 0; k < n; ++k) {
        perm[k] = k;
        slot_node[k] = 0;
    }
    node_start[0] = 0;
    node_end[0] = n;
    node_depth[0] = 0;
    levels[0] = 0;
    levels[1] = 1;
    i64 n_levels = 1, lo = 0, hi = 1;
    sizes[2] = 0;
    for (;;) {
        i64 count = hi;
        for (i64 node = lo; node < hi; ++node) {
            if (node_end[node] - node_start[node] > leaf_capacity && node_depth[node] < max_depth) {
                if (count + 4 > capacity) {
                    sizes[2] = 1;
                    goto done;
                }
                node_child[node] = count;
                const double half = 0.5 * node_half[node];
                for (int c = 0; c < 4; ++c) {
                    node_half[count + c] = half;
                    node_depth[count + c] = node_depth[node] + 1;
                    cursor[count + c] = 0;
                }
                const double cx = node_center[2 * node], cy = node_center[2 * node + 1];
                node_center[2 * count + 0] = cx - half;
                node_center[2 * count + 1] = cy + half;
                node_center[2 * (count + 1) + 0] = cx + half;
                node_center[2 * (count + 1) + 1] = cy + half;
                node_center[2 * (count + 2) + 0] = cx - half;
                node_center[2 * (count + 2) + 1] = cy - half;
                node_center[2 * (count + 3) + 0] = cx + half;
                node_center[2 * (count + 3) + 1] = cy - half;
                count += 4;
            } else {
                node_child[node] = -1;
            }
        }
        if (count == hi) break;
        for (i64 k = 0; k < n; ++k) {
            const i64 node = slot_node[k];
            const i64 child = node_child[node];
            if (child < 0) {
                quadrant[k] = -1;
                continue;
            }
            i64 q = 0;
            if (sorted_pos[k] > node_center[2 * node]) q += 1;
            if (sorted_pos[ld + k] < node_center[2 * node + 1]) q += 2;
            quadrant[k] = q;
            cursor[child + q] += 1;
        }
        for (i64 node = lo; node < hi; ++node) {
            const i64 child = node_child[node];
            if (child < 0) continue;
            i64 start = node_start[node];
            for (int q = 0; q < 4; ++q) {
                const i64 size = cursor[child + q];
                node_start[child + q] = start;
                node_end[child + q] = start + size;
                cursor[child + q] = start;
                start += size;
            }
        }
        for (i64 k = 0; k < n; ++k) {
            const i64 q = quadrant[k];
            i64 slot;
            if (q < 0) {
                slot = k;
                new_slot_node[slot] = slot_node[k];
            } else {
                const i64 child = node_child[slot_node[k]] + q;
                slot = cursor[child];
                cursor[child] += 1;
                new_slot_node[slot] = child;
            }
            new_perm[slot] = perm[k];
            new_pos[slot] = sorted_pos[k];
            new_pos[n + slot] = sorted_pos[ld + k];
        }
        for (i64 k = 0; k < n; ++k) {
            perm[k] = new_perm[k];
            slot_node[k] = new_slot_node[k];
            sorted_pos[k] = new_pos[k];
            sorted_pos[ld + k] = new_pos[n + k];
        }
        lo = hi;
        hi = count;
        n_levels += 1;
        levels[n_levels] = hi;
    }
    sizes[0] = hi;
    sizes[1] = n_levels;
done:
    free(slot_node);
    free(quadrant);
    free(cursor);
    free(new_perm);
    free(new_slot_node);
    free(new_pos);
}

/* ------------------------------------------------------------------------------------ */
/* quadtree_moments_upward/quadtree_moments_upward_pyccel.py:13-83                       */
/* ------------------------------------------------------------------------------------ */
void oracle_quadtree_moments_upward(const double *pos, i64 ldp, const double *charge,
                                    const i64 *perm, const i64 *levels, i64 n_levels,
                                    const double *node_center, const i64 *node_start,
                                    const i64 *node_end, const i64 *node_child, double *node_q,
                                    double *node_qabs, double *node_p, double *node_s) {
    for (i64 level = n_levels - 1; level >= 0; --level) {
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 node = levels[level]; node < levels[level + 1]; ++node) {
            double q = 0, qabs = 0, px = 0, py = 0, sxx = 0, sxy = 0, syy = 0;
            const double xmid = node_center[2 * node], ymid = node_center[2 * node + 1];
            const i64 first = node_child[node];
            if (first < 0) {
                for (i64 k = node_start[node]; k < node_end[node]; ++k) {
                    const i64 j = perm[k];
                    const double c = charge[j];
                    const double dx = pos[j] - xmid, dy = pos[ldp + j] - ymid;
                    q += c;
                    qabs += fabs(c);
                    px += c * dx;
                    py += c * dy;
                    sxx += c * dx * dx;
                    sxy += c * dx * dy;
                    syy += c * dy * dy;
                }
            } else {
                for (i64 child = first; child < first + 4; ++child) {
                    const double dx = node_center[2 * child] - xmid;
                    const double dy = node_center[2 * child + 1] - ymid;
                    const double cq = node_q[child];
                    const double cpx = node_p[2 * child], cpy = node_p[2 * child + 1];
                    q += cq;
                    qabs += node_qabs[child];
                    px += cpx + cq * dx;
                    py += cpy + cq * dy;
                    sxx += node_s[3 * child] + 2.0 * cpx * dx + cq * dx * dx;
                    sxy += node_s[3 * child + 1] + cpx * dy + cpy * dx + cq * dx * dy;
                    syy += node_s[3 * child + 2] + 2.0 * cpy * dy + cq * dy * dy;
                }
            }
            node_q[node] = q;
            node_qabs[node] = qabs;
            node_p[2 * node] = px;
            node_p[2 * node + 1] = py;
            node_s[3 * node] = sxx;
            node_s[3 * node + 1] = sxy;
            node_s[3 * node + 2] = syy;
        }
    }
}

/* flat_tree.py:224-231 (_binomial_table): C(n, k) for n, k < size, row-major. */
void oracle_binomial_table(i64 size, double *table) {
    memset(table, 0, sizeof(double) * (size_t)(size * size));
    for (i64 row = 0; row < size; ++row) {
        table[row * size] = 1.0;
        for (i64 col = 1; col <= row; ++col)
            table[row * size + col] = table[(row - 1) * size + col - 1] + table[(row - 1) * size + col];
    }
}

/* flat_tree.py:234-245 (_m2l_coefficients): coeff[k, l], (order+1)^2 row-major. */
void oracle_m2l_coefficients(i64 order, int symmetric, double *coeff) {
    const i64 size = 2 * order + 1;
    double *binomial = (double *)malloc(sizeof(double) * (size_t)(size * size));
    oracle_binomial_table(size, binomial);
    const i64 p1 = order + 1;
    memset(coeff, 0, sizeof(double) * (size_t)(p1 * p1));
    for (i64 m = 1; m <= order; ++m) coeff[m] = 1.0 / (double)m;
    for (i64 k = 1; k <= order; ++k) {
        coeff[k * p1] = 1.0 / (double)k;
        for (i64 m = 1; m <= order; ++m)
            if (k <= (symmetric ? order - m : order))
                coeff[k * p1 + m] = binomial[(k + m - 1) * size + m] / (double)k;
    }
    free(binomial);
}

/* ------------------------------------------------------------------------------------ */
/* log_moments_upward/log_moments_upward_pyccel.py:14-93.  binomial is (bsz, bsz).      */
/* ------------------------------------------------------------------------------------ */
void oracle_log_moments_upward(const double *pos, i64 ldp, const double *charge, const i64 *perm,
                               const i64 *levels, i64 n_levels, const double *node_center,
                               const i64 *node_start, const i64 *node_end, const i64 *node_child,
                               const double *binomial, i64 bsz, i64 order, double *node_qabs,
                               double *moment_re, double *moment_im) {
    const i64 p1 = order + 1;
    for (i64 level = n_levels - 1; level >= 0; --level) {
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 node = levels[level]; node < levels[level + 1]; ++node) {
            double sh_re[130], sh_im[130];
            const double xmid = node_center[2 * node], ymid = node_center[2 * node + 1];
            double *mre = moment_re + node * p1, *mim = moment_im + node * p1;
            for (i64 n = 0; n <= order; ++n) mre[n] = mim[n] = 0.0;
            const i64 first = node_child[node];
            double qabs = 0.0;
            if (first < 0) {
                for (i64 k = node_start[node]; k < node_end[node]; ++k) {
                    const i64 j = perm[k];
                    const double q = charge[j];
                    qabs += fabs(q);
                    const double dx = pos[j] - xmid, dy = pos[ldp + j] - ymid;
                    double pr = q, pi = 0.0;
                    for (i64 n = 0; n <= order; ++n) {
                        mre[n] += pr;
                        mim[n] += pi;
                        const double next_re = pr * dx - pi * dy;
                        pi = pr * dy + pi * dx;
                        pr = next_re;
                    }
                }
            } else {
                for (i64 child = first; child < first + 4; ++child) {
                    qabs += node_qabs[child];
                    if (node_end[child] == node_start[child]) continue;
                    const double s_re = node_center[2 * child] - xmid;
                    const double s_im = node_center[2 * child + 1] - ymid;
                    sh_re[0] = 1.0;
                    sh_im[0] = 0.0;
                    for (i64 n = 1; n <= order; ++n) {
                        sh_re[n] = sh_re[n - 1] * s_re - sh_im[n - 1] * s_im;
                        sh_im[n] = sh_re[n - 1] * s_im + sh_im[n - 1] * s_re;
                    }
                    const double *cre = moment_re + child * p1, *cim = moment_im + child * p1;
                    for (i64 n = 0; n <= order; ++n) {
                        double acc_re = 0.0, acc_im = 0.0;
                        for (i64 m = 0; m <= n; ++m) {
                            const double c = binomial[n * bsz + m];
                            const double a_re = cre[m] * c, a_im = cim[m] * c;
                            acc_re += a_re * sh_re[n - m] - a_im * sh_im[n - m];
                            acc_im += a_re * sh_im[n - m] + a_im * sh_re[n - m];
                        }
                        mre[n] += acc_re;
                        mim[n] += acc_im;
                    }
                }
            }
            node_qabs[node] = qabs;
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* evaluate_groups/evaluate_groups_pyccel.py:16-297: the Barnes-Hut walk by groups of   */
/* at most 62 targets with 64-bit lane masks.  spos (2, lds) and scharge are the sources */
/* in tree order, sindex = perm.  moment arrays are (C, order+1) (unused for Coulomb).   */
/* ------------------------------------------------------------------------------------ */
void oracle_evaluate_groups(const double *targets, i64 ldt, const i64 *exclude,
                            const i64 *group_start, i64 n_groups, int log_kernel,
                            int with_potential, const double *spos, i64 lds,
                            const double *scharge, const i64 *sindex, const double *node_center,
                            const double *node_half, const i64 *node_start, const i64 *node_end,
                            const i64 *node_child, const double *node_q, const double *node_qabs,
                            const double *node_p, const double *node_s, const double *moment_re,
                            const double *moment_im, i64 order, double theta, double softening,
                            i64 stack_size, double *phi, double *field, i64 ldf) {
    const double eps2 = softening * softening;
    const i64 p1 = order + 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (i64 g = 0; g < n_groups; ++g) {
        i64 *stack_node = (i64 *)malloc(sizeof(i64) * (size_t)stack_size);
        uint64_t *stack_mask = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)stack_size);
        double pot[64] = {0}, ex[64] = {0}, ey[64] = {0};
        i64 bad[64] = {0};
        double gx_t[64], gy_t[64];
        i64 gskip[64];
        const i64 t0 = group_start[g];
        const i64 nt = group_start[g + 1] - t0;
        if (nt <= 0) {
            free(stack_node);
            free(stack_mask);
            continue;
        }
        for (i64 b = 0; b < nt; ++b) {
            gx_t[b] = targets[t0 + b];
            gy_t[b] = targets[ldt + t0 + b];
            gskip[b] = exclude[t0 + b];
        }
        double xlo = gx_t[0], xhi = xlo, ylo = gy_t[0], yhi = ylo;
        for (i64 b = 1; b < nt; ++b) {
            xlo = fmin(xlo, gx_t[b]);
            xhi = fmax(xhi, gx_t[b]);
            ylo = fmin(ylo, gy_t[b]);
            yhi = fmax(yhi, gy_t[b]);
        }
        const uint64_t full = (nt >= 64) ? ~0ULL : ((1ULL << nt) - 1ULL);
        stack_node[0] = 0;
        stack_mask[0] = full;
        i64 top = 1;
        while (top > 0) {
            top -= 1;
            const i64 node = stack_node[top];
            const uint64_t mask = stack_mask[top];
            if (node_qabs[node] == 0.0) continue;
            const i64 first = node_child[node];
            if (first < 0) {
                if (mask == full) {
                    for (i64 k = node_start[node]; k < node_end[node]; ++k) {
                        const double c = scharge[k];
                        if (c == 0.0) continue;
                        const double sx = spos[k], sy = spos[lds + k];
                        const i64 sid = sindex[k];
                        for (i64 b = 0; b < nt; ++b) {
                            const double dx = gx_t[b] - sx, dy = gy_t[b] - sy;
                            double r2 = dx * dx + dy * dy + eps2;
                            const double cb = gskip[b] == sid ? 0.0 : c;
                            r2 = gskip[b] == sid ? 1.0 : r2;
                            bad[b] += r2 == 0.0 ? 1 : 0;
                            if (log_kernel) {
                                if (with_potential) pot[b] -= cb * log(r2);
                                ex[b] += 2.0 * cb * dx / r2;
                                ey[b] += 2.0 * cb * dy / r2;
                            } else {
                                const double inv_r = 1.0 / sqrt(r2);
                                const double inv_r3 = inv_r * inv_r * inv_r;
                                pot[b] += cb * inv_r;
                                ex[b] += cb * inv_r3 * dx;
                                ey[b] += cb * inv_r3 * dy;
                            }
                        }
                    }
                    continue;
                }
                for (i64 b = 0; b < nt; ++b) {
                    if (((mask >> b) & 1ULL) == 0 || bad[b] != 0) continue;
                    const double tx = gx_t[b], ty = gy_t[b];
                    const i64 skip = gskip[b];
                    for (i64 k = node_start[node]; k < node_end[node]; ++k) {
                        const double c = scharge[k];
                        if (sindex[k] == skip || c == 0.0) continue;
                        const double dx = tx - spos[k], dy = ty - spos[lds + k];
                        const double r2 = dx * dx + dy * dy + eps2;
                        if (r2 == 0.0) {
                            bad[b] = 1;
                            break;
                        }
                        if (log_kernel) {
                            if (with_potential) pot[b] -= c * log(r2);
                            ex[b] += 2.0 * c * dx / r2;
                            ey[b] += 2.0 * c * dy / r2;
                        } else {
                            const double inv_r = 1.0 / sqrt(r2);
                            const double inv_r3 = inv_r * inv_r * inv_r;
                            pot[b] += c * inv_r;
                            ex[b] += c * inv_r3 * dx;
                            ey[b] += c * inv_r3 * dy;
                        }
                    }
                }
                continue;
            }
            const double xmid = node_center[2 * node], ymid = node_center[2 * node + 1];
            const double half = node_half[node];
            double gx = fmax(fmax(xlo - xmid, xmid - xhi), 0.0);
            double gy = fmax(fmax(ylo - ymid, ymid - yhi), 0.0);
            const double dmin = sqrt(gx * gx + gy * gy);
            gx = fmax(fabs(xlo - xmid), fabs(xhi - xmid));
            gy = fmax(fabs(ylo - ymid), fabs(yhi - ymid));
            const double dmax = sqrt(gx * gx + gy * gy);
            uint64_t accept, reject;
            if ((xhi < xmid - half || xlo > xmid + half || yhi < ymid - half || ylo > ymid + half) &&
                2.0 * half < theta * dmin * (1.0 - 1.0e-12)) {
                accept = mask;
                reject = 0;
            } else if (2.0 * half >= theta * dmax * (1.0 + 1.0e-12)) {
                accept = 0;
                reject = mask;
            } else {
                accept = 0;
                reject = 0;
                for (i64 b = 0; b < nt; ++b) {
                    if (((mask >> b) & 1ULL) == 0) continue;
                    const double tx = targets[t0 + b], ty = targets[ldt + t0 + b];
                    const double rx = tx - xmid, ry = ty - ymid;
                    const int inside = tx >= xmid - half && tx <= xmid + half && ty >= ymid - half &&
                                       ty <= ymid + half;
                    if (log_kernel) {
                        const double d2 = rx * rx + ry * ry;
                        if (!inside && 4.0 * half * half < theta * theta * d2)
                            accept |= 1ULL << b;
                        else
                            reject |= 1ULL << b;
                    } else {
                        const double distance = sqrt(rx * rx + ry * ry);
                        if (!inside && 2.0 * half < theta * distance)
                            accept |= 1ULL << b;
                        else
                            reject |= 1ULL << b;
                    }
                }
            }
            if (accept != 0) {
                for (i64 b = 0; b < nt; ++b) {
                    if (((accept >> b) & 1ULL) == 0 || bad[b] != 0) continue;
                    const double rx = targets[t0 + b] - xmid, ry = targets[ldt + t0 + b] - ymid;
                    if (log_kernel) {
                        const double d2 = rx * rx + ry * ry;
                        const double w_re = rx / d2, w_im = -ry / d2;
                        double prev_re = 1.0, prev_im = 0.0, cur_re = w_re, cur_im = w_im;
                        double f_re = 0.0, f_im = 0.0, series = 0.0;
                        const double *are = moment_re + node * p1, *aim = moment_im + node * p1;
                        for (i64 n = 0; n <= order; ++n) {
                            const double a_re = are[n], a_im = aim[n];
                            f_re += a_re * cur_re - a_im * cur_im;
                            f_im += a_re * cur_im + a_im * cur_re;
                            if (n > 0) series += (a_re * prev_re - a_im * prev_im) / (double)n;
                            prev_re = cur_re;
                            prev_im = cur_im;
                            const double next_re = cur_re * w_re - cur_im * w_im;
                            cur_im = cur_re * w_im + cur_im * w_re;
                            cur_re = next_re;
                        }
                        if (with_potential) pot[b] += -are[0] * log(d2) + 2.0 * series;
                        ex[b] += 2.0 * f_re;
                        ey[b] -= 2.0 * f_im;
                    } else {
                        const double inv_r = 1.0 / sqrt(rx * rx + ry * ry + eps2);
                        const double inv_r2 = inv_r * inv_r;
                        const double inv_r3 = inv_r2 * inv_r;
                        const double inv_r5 = inv_r3 * inv_r2;
                        const double inv_r7 = inv_r5 * inv_r2;
                        const double q = node_q[node];
                        const double px = node_p[2 * node], py = node_p[2 * node + 1];
                        const double sxx = node_s[3 * node], sxy = node_s[3 * node + 1],
                                     syy = node_s[3 * node + 2];
                        const double pr = px * rx + py * ry;
                        const double srx = sxx * rx + sxy * ry;
                        const double sry = sxy * rx + syy * ry;
                        const double rsr = rx * srx + ry * sry;
                        const double trace = sxx + syy;
                        pot[b] += q * inv_r + pr * inv_r3 + 0.5 * (3.0 * rsr * inv_r5 - trace * inv_r3);
                        const double radial =
                            q * inv_r3 + 3.0 * pr * inv_r5 + 7.5 * rsr * inv_r7 - 1.5 * trace * inv_r5;
                        ex[b] += radial * rx - px * inv_r3 - 3.0 * srx * inv_r5;
                        ey[b] += radial * ry - py * inv_r3 - 3.0 * sry * inv_r5;
                    }
                }
            }
            if (reject != 0) {
                for (int child = 0; child < 4; ++child) {
                    stack_node[top] = first + child;
                    stack_mask[top] = reject;
                    top += 1;
                }
            }
        }
        for (i64 b = 0; b < nt; ++b) {
            const i64 i = t0 + b;
            if (bad[b] != 0) {
                phi[i] = TD3_BAD;
            } else {
                phi[i] = pot[b];
                field[i] = ex[b];
                field[ldf + i] = ey[b];
            }
        }
        free(stack_node);
        free(stack_mask);
    }
}

/* ------------------------------------------------------------------------------------ */
/* fmm_traverse/fmm_traverse_pyccel.py:15-106: one generation of the mutual traversal.  */
/* frontier, next_frontier, m2l, p2p are (cap, 2); counts[4].                            */
/* ------------------------------------------------------------------------------------ */
void oracle_fmm_traverse(const double *node_center, const double *node_half, const i64 *node_start,
                         const i64 *node_end, const i64 *node_child, double theta,
                         const i64 *frontier, i64 n_frontier, i64 *next_frontier, i64 cap_next,
                         i64 *m2l, i64 cap_m2l, i64 *p2p, i64 cap_p2p, i64 *counts) {
    const double root2 = sqrt(2.0);
    for (i64 k = 0; k < n_frontier; ++k) {
        const i64 a = frontier[2 * k], b = frontier[2 * k + 1];
        if (node_end[a] == node_start[a] || node_end[b] == node_start[b]) continue;
        const int leaf_a = node_child[a] < 0, leaf_b = node_child[b] < 0;
        if (a == b) {
            if (leaf_a) {
                const i64 i = counts[1];
                counts[1] = i + 1;
                if (i < cap_p2p) {
                    p2p[2 * i] = a;
                    p2p[2 * i + 1] = a;
                } else {
                    counts[3] = -1;
                }
                continue;
            }
            const i64 first = node_child[a];
            i64 i = counts[2];
            counts[2] = i + 10;
            if (i + 10 > cap_next) {
                counts[3] = -1;
                continue;
            }
            for (int c = 0; c < 4; ++c)
                for (int d = c; d < 4; ++d) {
                    next_frontier[2 * i] = first + c;
                    next_frontier[2 * i + 1] = first + d;
                    i += 1;
                }
            continue;
        }
        const double dx = node_center[2 * a] - node_center[2 * b];
        const double dy = node_center[2 * a + 1] - node_center[2 * b + 1];
        const double distance = sqrt(dx * dx + dy * dy);
        const double radii = root2 * (node_half[a] + node_half[b]);
        if (radii < theta * distance) {
            const i64 i = counts[0];
            counts[0] = i + 1;
            if (i < cap_m2l) {
                m2l[2 * i] = a;
                m2l[2 * i + 1] = b;
            } else {
                counts[3] = -1;
            }
            continue;
        }
        if (leaf_a && leaf_b) {
            const i64 i = counts[1];
            counts[1] = i + 1;
            if (i < cap_p2p) {
                p2p[2 * i] = a;
                p2p[2 * i + 1] = b;
            } else {
                counts[3] = -1;
            }
            continue;
        }
        const i64 i = counts[2];
        counts[2] = i + 4;
        if (i + 4 > cap_next) {
            counts[3] = -1;
            continue;
        }
        if (leaf_b || (!leaf_a && node_half[a] >= node_half[b])) {
            const i64 first = node_child[a];
            for (int c = 0; c < 4; ++c) {
                next_frontier[2 * (i + c)] = first + c;
                next_frontier[2 * (i + c) + 1] = b;
            }
        } else {
            const i64 first = node_child[b];
            for (int c = 0; c < 4; ++c) {
                next_frontier[2 * (i + c)] = a;
                next_frontier[2 * (i + c) + 1] = first + c;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* pair_csr/pair_csr_pyccel.py:12-45.  row_start (n_nodes + 1) must be zero on entry.   */
/* ------------------------------------------------------------------------------------ */
void oracle_pair_csr(const i64 *pairs, i64 count, int mutual, i64 n_nodes, i64 *row_start,
                     i64 *partner) {
    i64 *cursor = (i64 *)malloc(sizeof(i64) * (size_t)(n_nodes > 0 ? n_nodes : 1));
    for (i64 p = 0; p < count; ++p) {
        const i64 a = pairs[2 * p], b = pairs[2 * p + 1];
        row_start[a + 1] += 1;
        if (mutual && a != b) row_start[b + 1] += 1;
    }
    for (i64 n = 0; n < n_nodes; ++n) {
        row_start[n + 1] += row_start[n];
        cursor[n] = row_start[n];
    }
    for (i64 p = 0; p < count; ++p) {
        const i64 a = pairs[2 * p], b = pairs[2 * p + 1];
        partner[cursor[a]++] = b;
        if (mutual && a != b) partner[cursor[b]++] = a;
    }
    free(cursor);
}

/* ------------------------------------------------------------------------------------ */
/* fmm_m2l_host/fmm_m2l_host_pyccel.py:16-105: M2L by destination row, real matrix      */
/* coeff[k, l] applied to the scaled moments a_k w^k.  Arrays (C, order+1).              */
/* ------------------------------------------------------------------------------------ */
void oracle_fmm_m2l_host(const i64 *row_start, const i64 *sources, i64 n_nodes,
                         const double *dst_center, const double *src_center,
                         const double *moment_re, const double *moment_im, const double *coeff,
                         i64 order, double *local_re, double *local_im) {
    const i64 p1 = order + 1;
    const i64 n_chunks = (n_nodes + 63) / 64;
#pragma omp parallel for schedule(dynamic, 1)
    for (i64 chunk = 0; chunk < n_chunks; ++chunk) {
        double b_re[65], b_im[65], s_re[65], s_im[65], pw_re[65], pw_im[65];
        const i64 last = (64 * chunk + 64 < n_nodes) ? 64 * chunk + 64 : n_nodes;
        for (i64 dst = 64 * chunk; dst < last; ++dst) {
            if (row_start[dst + 1] == row_start[dst]) continue;
            for (i64 l = 0; l <= order; ++l) b_re[l] = b_im[l] = 0.0;
            const double cx = dst_center[2 * dst], cy = dst_center[2 * dst + 1];
            for (i64 e = row_start[dst]; e < row_start[dst + 1]; ++e) {
                const i64 src = sources[e];
                const double t_re = cx - src_center[2 * src], t_im = cy - src_center[2 * src + 1];
                const double t2 = t_re * t_re + t_im * t_im;
                const double w_re = t_re / t2, w_im = -t_im / t2;
                pw_re[0] = 1.0;
                pw_im[0] = 0.0;
                for (i64 k = 1; k <= order; ++k) {
                    pw_re[k] = pw_re[k - 1] * w_re - pw_im[k - 1] * w_im;
                    pw_im[k] = pw_re[k - 1] * w_im + pw_im[k - 1] * w_re;
                }
                for (i64 l = 0; l <= order; ++l) s_re[l] = s_im[l] = 0.0;
                for (i64 k = 0; k <= order; ++k) {
                    const double a_re = moment_re[src * p1 + k], a_im = moment_im[src * p1 + k];
                    const double ar = a_re * pw_re[k] - a_im * pw_im[k];
                    const double ai = a_re * pw_im[k] + a_im * pw_re[k];
                    for (i64 l = 0; l <= order; ++l) {
                        const double c = coeff[k * p1 + l];
                        s_re[l] += c * ar;
                        s_im[l] += c * ai;
                    }
                }
                const double a_re = moment_re[src * p1], a_im = moment_im[src * p1];
                const double log_re = 0.5 * log(t2), log_im = atan2(t_im, t_re);
                b_re[0] += a_re * log_re - a_im * log_im - s_re[0];
                b_im[0] += a_re * log_im + a_im * log_re - s_im[0];
                double sign = 1.0;
                for (i64 l = 1; l <= order; ++l) {
                    b_re[l] += sign * (s_re[l] * pw_re[l] - s_im[l] * pw_im[l]);
                    b_im[l] += sign * (s_re[l] * pw_im[l] + s_im[l] * pw_re[l]);
                    sign = -sign;
                }
            }
            for (i64 l = 0; l <= order; ++l) {
                local_re[dst * p1 + l] += b_re[l];
                local_im[dst * p1 + l] += b_im[l];
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* fmm_downward/fmm_downward_pyccel.py:15-64: L2L from nodes [first, last) to children. */
/* ------------------------------------------------------------------------------------ */
void oracle_fmm_downward(i64 first_node, i64 last_node, const double *node_center,
                         const i64 *node_child, const double *binomial, i64 bsz, i64 order,
                         double *local_re, double *local_im) {
    const i64 p1 = order + 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 node = first_node; node < last_node; ++node) {
        double sh_re[130], sh_im[130];
        const i64 first = node_child[node];
        if (first < 0) continue;
        for (i64 child = first; child < first + 4; ++child) {
            const double s_re = node_center[2 * child] - node_center[2 * node];
            const double s_im = node_center[2 * child + 1] - node_center[2 * node + 1];
            sh_re[0] = 1.0;
            sh_im[0] = 0.0;
            for (i64 k = 1; k <= order; ++k) {
                sh_re[k] = sh_re[k - 1] * s_re - sh_im[k - 1] * s_im;
                sh_im[k] = sh_re[k - 1] * s_im + sh_im[k - 1] * s_re;
            }
            for (i64 m = 0; m <= order; ++m) {
                double acc_re = 0.0, acc_im = 0.0;
                for (i64 l = m; l <= order; ++l) {
                    const double c = binomial[l * bsz + m];
                    const double b_re = local_re[node * p1 + l] * c;
                    const double b_im = local_im[node * p1 + l] * c;
                    acc_re += b_re * sh_re[l - m] - b_im * sh_im[l - m];
                    acc_im += b_re * sh_im[l - m] + b_im * sh_re[l - m];
                }
                local_re[child * p1 + m] += acc_re;
                local_im[child * p1 + m] += acc_im;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* fmm_evaluate_host/fmm_evaluate_host_pyccel.py:14-100: L2P (parallel) then mutual P2P */
/* (serial) in tree order.  spos (2, lds); phi (N), field (2, ldf) zero on entry.        */
/* ------------------------------------------------------------------------------------ */
void oracle_fmm_evaluate_host(const double *spos, i64 lds, const double *scharge, i64 node_count,
                              const double *node_center, const i64 *node_start,
                              const i64 *node_end, const i64 *node_child, const double *local_re,
                              const double *local_im, i64 order, const i64 *pairs, i64 count,
                              double softening, int with_potential, double *phi, double *field,
                              i64 ldf, i64 *status) {
    const i64 p1 = order + 1;
    const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 node = 0; node < node_count; ++node) {
        if (node_child[node] >= 0) continue;
        const double *bre = local_re + node * p1, *bim = local_im + node * p1;
        for (i64 k = node_start[node]; k < node_end[node]; ++k) {
            const double u_re = spos[k] - node_center[2 * node];
            const double u_im = spos[lds + k] - node_center[2 * node + 1];
            double w_re = bre[order], w_im = bim[order], d_re = 0.0, d_im = 0.0;
            for (i64 l = order - 1; l >= 0; --l) {
                const double nd_re = d_re * u_re - d_im * u_im + w_re;
                d_im = d_re * u_im + d_im * u_re + w_im;
                d_re = nd_re;
                const double nw_re = w_re * u_re - w_im * u_im + bre[l];
                w_im = w_re * u_im + w_im * u_re + bim[l];
                w_re = nw_re;
            }
            phi[k] += -2.0 * w_re;
            field[k] += 2.0 * d_re;
            field[ldf + k] -= 2.0 * d_im;
        }
    }
    for (i64 pair = 0; pair < count; ++pair) {
        const i64 a = pairs[2 * pair], b = pairs[2 * pair + 1];
        for (i64 ka = node_start[a]; ka < node_end[a]; ++ka) {
            const double xi = spos[ka], yi = spos[lds + ka], qi = scharge[ka];
            double pi = 0.0, fxi = 0.0, fyi = 0.0;
            const i64 start_b = (a == b) ? ka + 1 : node_start[b];
            for (i64 kb = start_b; kb < node_end[b]; ++kb) {
                const double qj = scharge[kb];
                const double dx = xi - spos[kb], dy = yi - spos[lds + kb];
                const double r2 = dx * dx + dy * dy + eps2;
                if (r2 == 0.0) {
                    if (qi != 0.0 && qj != 0.0) {
                        status[0] = ka + 1;
                        return;
                    }
                    continue;
                }
                if (with_potential) {
                    const double log_r2 = log(r2);
                    pi -= qj * log_r2;
                    phi[kb] -= qi * log_r2;
                }
                const double inv = 2.0 / r2;
                const double fx = dx * inv, fy = dy * inv;
                fxi += qj * fx;
                fyi += qj * fy;
                field[kb] -= qi * fx;
                field[ldf + kb] -= qi * fy;
            }
            phi[ka] += pi;
            field[ka] += fxi;
            field[ldf + ka] += fyi;
        }
    }
}

/* ------------------------------------------------------------------------------------ */
/* dual_traverse/dual_traverse_pyccel.py:14-86: depth-first target x source traversal,  */
/* one way.  stack (cap_stack, 2); counts[3] = {m2l, p2p, overflow}.                     */
/* ------------------------------------------------------------------------------------ */
void oracle_dual_traverse(const double *t_center, const double *t_half, const i64 *t_start,
                          const i64 *t_end, const i64 *t_child, const double *s_center,
                          const double *s_half, const i64 *s_start, const i64 *s_end,
                          const i64 *s_child, double theta, i64 *stack, i64 cap_stack, i64 *m2l,
                          i64 cap_m2l, i64 *p2p, i64 cap_p2p, i64 *counts) {
    const double root2 = sqrt(2.0);
    counts[0] = counts[1] = counts[2] = 0;
    stack[0] = 0;
    stack[1] = 0;
    i64 top = 1;
    while (top > 0) {
        top -= 1;
        const i64 a = stack[2 * top], b = stack[2 * top + 1];
        if (t_end[a] == t_start[a] || s_end[b] == s_start[b]) continue;
        const double dx = t_center[2 * a] - s_center[2 * b];
        const double dy = t_center[2 * a + 1] - s_center[2 * b + 1];
        if (root2 * (t_half[a] + s_half[b]) < theta * sqrt(dx * dx + dy * dy)) {
            const i64 i = counts[0]++;
            if (i < cap_m2l) {
                m2l[2 * i] = a;
                m2l[2 * i + 1] = b;
            } else {
                counts[2] = 1;
            }
            continue;
        }
        const int leaf_a = t_child[a] < 0, leaf_b = s_child[b] < 0;
        if (leaf_a && leaf_b) {
            const i64 i = counts[1]++;
            if (i < cap_p2p) {
                p2p[2 * i] = a;
                p2p[2 * i + 1] = b;
            } else {
                counts[2] = 1;
            }
            continue;
        }
        if (top + 4 > cap_stack) {
            counts[2] = 1;
            return;
        }
        if (leaf_b || (!leaf_a && t_half[a] >= s_half[b])) {
            for (int c = 0; c < 4; ++c) {
                stack[2 * (top + c)] = t_child[a] + c;
                stack[2 * (top + c) + 1] = b;
            }
        } else {
            for (int c = 0; c < 4; ++c) {
                stack[2 * (top + c)] = a;
                stack[2 * (top + c) + 1] = s_child[b] + c;
            }
        }
        top += 4;
    }
}

/* ------------------------------------------------------------------------------------ */
/* log_targets_host/log_targets_host_pyccel.py:14-94: per target leaf, L2P + exact sums  */
/* over its source leaves.  tpos (2, ldt), spos (2, lds); field (2, ldf).                 */
/* ------------------------------------------------------------------------------------ */
void oracle_log_targets_host(const double *tpos, i64 ldt, const i64 *texclude, i64 t_count,
                             const double *t_center, const i64 *t_start, const i64 *t_end,
                             const i64 *t_child, const double *local_re, const double *local_im,
                             i64 order, const i64 *row_start, const i64 *partners,
                             const double *spos, i64 lds, const double *scharge, const i64 *sindex,
                             const i64 *s_start, const i64 *s_end, double softening,
                             int with_potential, double *phi, double *field, i64 ldf) {
    const i64 p1 = order + 1;
    const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 node = 0; node < t_count; ++node) {
        if (t_child[node] >= 0) continue;
        const double *bre = local_re + node * p1, *bim = local_im + node * p1;
        for (i64 k = t_start[node]; k < t_end[node]; ++k) {
            const double u_re = tpos[k] - t_center[2 * node];
            const double u_im = tpos[ldt + k] - t_center[2 * node + 1];
            double w_re = bre[order], w_im = bim[order], d_re = 0.0, d_im = 0.0;
            for (i64 l = order - 1; l >= 0; --l) {
                const double nd_re = d_re * u_re - d_im * u_im + w_re;
                d_im = d_re * u_im + d_im * u_re + w_im;
                d_re = nd_re;
                const double nw_re = w_re * u_re - w_im * u_im + bre[l];
                w_im = w_re * u_im + w_im * u_re + bim[l];
                w_re = nw_re;
            }
            double pot = -2.0 * w_re, ex = 2.0 * d_re, ey = -2.0 * d_im;
            int bad = 0;
            const i64 skip = texclude[k];
            for (i64 e = row_start[node]; e < row_start[node + 1]; ++e) {
                const i64 src = partners[e];
                for (i64 j = s_start[src]; j < s_end[src]; ++j) {
                    const double c = scharge[j];
                    if (c == 0.0 || sindex[j] == skip) continue;
                    const double dx = tpos[k] - spos[j], dy = tpos[ldt + k] - spos[lds + j];
                    const double r2 = dx * dx + dy * dy + eps2;
                    if (r2 == 0.0) {
                        bad = 1;
                        continue;
                    }
                    if (with_potential) pot -= c * log(r2);
                    ex += 2.0 * c * dx / r2;
                    ey += 2.0 * c * dy / r2;
                }
            }
            if (bad)
                phi[k] = TD3_BAD;
            else
                phi[k] = with_potential ? pot : 0.0;
            field[k] = ex;
            field[ldf + k] = ey;
        }
    }
}

/* ==================================================================================== */
/* The Coulomb (Cartesian) FMM: kernels/cart_*_pyccel.py with the tables of _cartesian.py */
/* (built by oracle.py).  nc = (p+1)(p+2)/2 coefficients per cell.                       */
/* ==================================================================================== */

/* cart_moments_upward/cart_moments_upward_pyccel.py:14-82 */
void oracle_cart_moments_upward(const double *pos, i64 ldp, const double *charge, const i64 *perm,
                                const i64 *levels, i64 n_levels, const double *node_center,
                                const i64 *node_start, const i64 *node_end, const i64 *node_child,
                                const i64 *mono_parent, const i64 *mono_dir, const i64 *shift,
                                const double *shift_coef, i64 n_shift, i64 nc, double *moments) {
    for (i64 level = n_levels - 1; level >= 0; --level) {
#pragma omp parallel for schedule(dynamic, 64)
        for (i64 node = levels[level]; node < levels[level + 1]; ++node) {
            double mono[2200];
            const double xmid = node_center[2 * node], ymid = node_center[2 * node + 1];
            double *m = moments + node * nc;
            for (i64 i = 0; i < nc; ++i) m[i] = 0.0;
            const i64 first = node_child[node];
            if (first < 0) {
                for (i64 k = node_start[node]; k < node_end[node]; ++k) {
                    const i64 j = perm[k];
                    mono[0] = charge[j];
                    for (i64 i = 1; i < nc; ++i)
                        mono[i] = mono[mono_parent[i]] *
                                  (mono_dir[i] == 0 ? pos[j] - xmid : pos[ldp + j] - ymid);
                    for (i64 i = 0; i < nc; ++i) m[i] += mono[i];
                }
            } else {
                for (i64 child = first; child < first + 4; ++child) {
                    if (node_end[child] == node_start[child]) continue;
                    mono[0] = 1.0;
                    for (i64 i = 1; i < nc; ++i)
                        mono[i] = mono[mono_parent[i]] * (mono_dir[i] == 0
                                                              ? node_center[2 * child] - xmid
                                                              : node_center[2 * child + 1] - ymid);
                    const double *cm = moments + child * nc;
                    for (i64 t = 0; t < n_shift; ++t)
                        m[shift[3 * t]] += shift_coef[t] * cm[shift[3 * t + 1]] * mono[shift[3 * t + 2]];
                }
            }
        }
    }
}

/* cart_m2l_host/cart_m2l_host_pyccel.py:16-80 */
void oracle_cart_m2l_host(const i64 *row_start, const i64 *sources, i64 n_nodes,
                          const double *dst_center, const double *src_center,
                          const double *moments, i64 nc, double softening, const i64 *ex,
                          const i64 *ey, const i64 *rec, const i64 *m2l, const double *m2l_coef,
                          i64 n_terms, double *local) {
    const double eps2 = softening * softening;
    const i64 n_chunks = (n_nodes + 63) / 64;
#pragma omp parallel for schedule(dynamic, 1)
    for (i64 chunk = 0; chunk < n_chunks; ++chunk) {
        double acc[2200], coef[2200];
        const i64 last = (64 * chunk + 64 < n_nodes) ? 64 * chunk + 64 : n_nodes;
        for (i64 dst = 64 * chunk; dst < last; ++dst) {
            if (row_start[dst + 1] == row_start[dst]) continue;
            for (i64 i = 0; i < nc; ++i) acc[i] = 0.0;
            const double cx = dst_center[2 * dst], cy = dst_center[2 * dst + 1];
            for (i64 e = row_start[dst]; e < row_start[dst + 1]; ++e) {
                const i64 src = sources[e];
                const double tx = cx - src_center[2 * src], ty = cy - src_center[2 * src + 1];
                const double r2 = tx * tx + ty * ty + eps2;
                coef[0] = 1.0 / sqrt(r2);
                for (i64 i = 1; i < nc; ++i) {
                    const i64 n = ex[i] + ey[i];
                    double s = 0.0;
                    if (rec[4 * i] >= 0) s += (2 * n - 1) * tx * coef[rec[4 * i]];
                    if (rec[4 * i + 1] >= 0) s += (2 * n - 1) * ty * coef[rec[4 * i + 1]];
                    if (rec[4 * i + 2] >= 0) s += (n - 1) * coef[rec[4 * i + 2]];
                    if (rec[4 * i + 3] >= 0) s += (n - 1) * coef[rec[4 * i + 3]];
                    coef[i] = -s / (n * r2);
                }
                const double *ms = moments + src * nc;
                for (i64 t = 0; t < n_terms; ++t)
                    acc[m2l[3 * t]] += m2l_coef[t] * ms[m2l[3 * t + 1]] * coef[m2l[3 * t + 2]];
            }
            for (i64 i = 0; i < nc; ++i) local[dst * nc + i] += acc[i];
        }
    }
}

/* cart_downward/cart_downward_pyccel.py:14-55 */
void oracle_cart_downward(const i64 *levels, i64 n_levels, const double *node_center,
                          const i64 *node_child, const i64 *mono_parent, const i64 *mono_dir,
                          const i64 *shift, const double *shift_coef, i64 n_shift, i64 nc,
                          double *local) {
    for (i64 level = 0; level < n_levels - 1; ++level) {
#pragma omp parallel for schedule(dynamic, 16)
        for (i64 node = levels[level]; node < levels[level + 1]; ++node) {
            double mono[2200];
            const i64 first = node_child[node];
            if (first < 0) continue;
            for (i64 child = first; child < first + 4; ++child) {
                mono[0] = 1.0;
                for (i64 i = 1; i < nc; ++i)
                    mono[i] = mono[mono_parent[i]] *
                              (mono_dir[i] == 0 ? node_center[2 * child] - node_center[2 * node]
                                                : node_center[2 * child + 1] - node_center[2 * node + 1]);
                for (i64 t = 0; t < n_shift; ++t)
                    local[child * nc + shift[3 * t + 1]] +=
                        shift_coef[t] * local[node * nc + shift[3 * t]] * mono[shift[3 * t + 2]];
            }
        }
    }
}

/* cart_evaluate_host/cart_evaluate_host_pyccel.py:16-110 (far part: local expansions at
 * the leaves' particles; near part: the mutual softened pairs, serial). */
void oracle_cart_evaluate_host(const double *spos, i64 lds, const double *scharge, i64 node_count,
                               const double *node_center, const i64 *node_start,
                               const i64 *node_end, const i64 *node_child, const double *local,
                               i64 nc, const i64 *mono_parent, const i64 *mono_dir, const i64 *ex,
                               const i64 *ey, const i64 *grad, const i64 *pairs, i64 count,
                               double softening, double *phi, double *field, i64 ldf, i64 *status) {
    const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 node = 0; node < node_count; ++node) {
        double mono[2200];
        if (node_child[node] >= 0) continue;
        const double *L = local + node * nc;
        for (i64 k = node_start[node]; k < node_end[node]; ++k) {
            mono[0] = 1.0;
            for (i64 i = 1; i < nc; ++i)
                mono[i] = mono[mono_parent[i]] * (mono_dir[i] == 0 ? spos[k] - node_center[2 * node]
                                                                   : spos[lds + k] - node_center[2 * node + 1]);
            double w = 0.0, gx = 0.0, gy = 0.0;
            for (i64 i = 0; i < nc; ++i) {
                w += L[i] * mono[i];
                if (grad[2 * i] >= 0) gx += ex[i] * L[i] * mono[grad[2 * i]];
                if (grad[2 * i + 1] >= 0) gy += ey[i] * L[i] * mono[grad[2 * i + 1]];
            }
            phi[k] += w;
            field[k] -= gx;
            field[ldf + k] -= gy;
        }
    }
    for (i64 pair = 0; pair < count; ++pair) {
        const i64 a = pairs[2 * pair], b = pairs[2 * pair + 1];
        for (i64 ka = node_start[a]; ka < node_end[a]; ++ka) {
            const double xi = spos[ka], yi = spos[lds + ka], qi = scharge[ka];
            double pi = 0.0, fxi = 0.0, fyi = 0.0;
            const i64 start_b = (a == b) ? ka + 1 : node_start[b];
            for (i64 kb = start_b; kb < node_end[b]; ++kb) {
                const double qj = scharge[kb];
                const double dx = xi - spos[kb], dy = yi - spos[lds + kb];
                const double r2 = dx * dx + dy * dy + eps2;
                if (r2 == 0.0) {
                    if (qi != 0.0 && qj != 0.0) {
                        status[0] = ka + 1;
                        return;
                    }
                    continue;
                }
                const double inv_r = 1.0 / sqrt(r2);
                const double inv_r3 = inv_r * inv_r * inv_r;
                pi += qj * inv_r;
                phi[kb] += qi * inv_r;
                fxi += qj * inv_r3 * dx;
                fyi += qj * inv_r3 * dy;
                field[kb] -= qi * inv_r3 * dx;
                field[ldf + kb] -= qi * inv_r3 * dy;
            }
            phi[ka] += pi;
            field[ka] += fxi;
            field[ldf + ka] += fyi;
        }
    }
}

/* cart_targets_host/cart_targets_host_pyccel.py:16-98 */
void oracle_cart_targets_host(const double *tpos, i64 ldt, const i64 *texclude, i64 t_count,
                              const double *t_center, const i64 *t_start, const i64 *t_end,
                              const i64 *t_child, const double *local, i64 nc,
                              const i64 *mono_parent, const i64 *mono_dir, const i64 *ex_pow,
                              const i64 *ey_pow, const i64 *grad, const i64 *row_start,
                              const i64 *partners, const double *spos, i64 lds,
                              const double *scharge, const i64 *sindex, const i64 *s_start,
                              const i64 *s_end, double softening, double *phi, double *field,
                              i64 ldf) {
    const double eps2 = softening * softening;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 node = 0; node < t_count; ++node) {
        double mono[2200];
        if (t_child[node] >= 0) continue;
        const double *L = local + node * nc;
        for (i64 k = t_start[node]; k < t_end[node]; ++k) {
            mono[0] = 1.0;
            for (i64 i = 1; i < nc; ++i)
                mono[i] = mono[mono_parent[i]] * (mono_dir[i] == 0 ? tpos[k] - t_center[2 * node]
                                                                   : tpos[ldt + k] - t_center[2 * node + 1]);
            double pot = 0.0, ex = 0.0, ey = 0.0;
            for (i64 i = 0; i < nc; ++i) {
                pot += L[i] * mono[i];
                if (grad[2 * i] >= 0) ex -= ex_pow[i] * L[i] * mono[grad[2 * i]];
                if (grad[2 * i + 1] >= 0) ey -= ey_pow[i] * L[i] * mono[grad[2 * i + 1]];
            }
            int bad = 0;
            const i64 skip = texclude[k];
            for (i64 e = row_start[node]; e < row_start[node + 1]; ++e) {
                const i64 src = partners[e];
                for (i64 j = s_start[src]; j < s_end[src]; ++j) {
                    const double c = scharge[j];
                    if (c == 0.0 || sindex[j] == skip) continue;
                    const double dx = tpos[k] - spos[j], dy = tpos[ldt + k] - spos[lds + j];
                    const double r2 = dx * dx + dy * dy + eps2;
                    if (r2 == 0.0) {
                        bad = 1;
                        continue;
                    }
                    const double inv_r = 1.0 / sqrt(r2);
                    const double inv_r3 = inv_r * inv_r * inv_r;
                    pot += c * inv_r;
                    ex += c * inv_r3 * dx;
                    ey += c * inv_r3 * dy;
                }
            }
            phi[k] = bad ? TD3_BAD : pot;
            field[k] = ex;
            field[ldf + k] = ey;
        }
    }
}
