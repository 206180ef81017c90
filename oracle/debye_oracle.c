/*
 * Test infrastructure: C restatement of the reference 'flat' CPU processor
 * arithmetic for F(Q) and its gradient, multithreaded with pthreads.  Used as
 * the timed CPU baseline and cross-checked against oracle/scatter.py.
 *
 * Follows (pyIID): kernels/cpu_flat.py:14-153 (float32 pair displacement,
 * distance, normalisation f_i f_j, omega = sin(Q r)/r, grad omega =
 * (Q cos(Q r) - omega)/r^2 d, products in float32), atomics/cpu_atomics.py:
 * 60-102 (per-chunk float64 F sums; float32 gradient scatter
 * grad[i] -= g, grad[j] += g, kernels/cpu_experimental.py:8-15) and
 * cpu_wrappers/flat_multi_cpu_wrap.py:22-102 (chunk sums, na = N *
 * mean(f_i f_j) in float32, F = 2 * sum / na, grad = sum / na, nan_to_num).
 * Q_m = float32(qbin) * m in float32 (cpu_flat.py:143-147).
 *
 * Each thread owns a contiguous range of the flat pair list (a "chunk" of the
 * reference) and private accumulators; `q_index` selects the Q bins to
 * evaluate (the cost is linear in their number).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  const float *pos, *scat;
  int64_t n, k0, k1;
  int nq, nqi;
  const int *q_index;
  float qbin;
  int want_grad;
  double *fsum;  /* nqi */
  float *grad;   /* n x 3 x nqi, or NULL */
} task_t;

static void k_to_ij(int64_t k, int64_t *i, int64_t *j) {
  int64_t ii = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)k)) / 2.0);
  if (ii * (ii - 1) / 2 > k) ii--;
  if ((ii + 1) * ii / 2 <= k) ii++;
  *i = ii;
  *j = k - ii * (ii - 1) / 2;
}

static void *worker(void *arg) {
  task_t *t = (task_t *)arg;
  const int nqi = t->nqi;
  float *sv = (float *)malloc(sizeof(float) * nqi);
  for (int a = 0; a < nqi; ++a) sv[a] = t->qbin * (float)t->q_index[a];
  int64_t i, j;
  k_to_ij(t->k0, &i, &j);
  for (int64_t k = t->k0; k < t->k1; ++k) {
    float d[3];
    for (int c = 0; c < 3; ++c) d[c] = t->pos[3 * i + c] - t->pos[3 * j + c];
    float tmp = 0.0f;
    for (int c = 0; c < 3; ++c) tmp += d[c] * d[c];
    const float r = sqrtf(tmp);
    const float *fi = t->scat + (size_t)i * t->nq;
    const float *fj = t->scat + (size_t)j * t->nq;
    for (int a = 0; a < nqi; ++a) {
      const int qx = t->q_index[a];
      const float norm = fi[qx] * fj[qx];
      const float om = sinf(sv[a] * r) / r;
      t->fsum[a] += (double)(norm * om);
      if (t->want_grad) {
        float g = sv[a] * cosf(sv[a] * r) - om;
        g /= r * r;
        float *gi = t->grad + ((size_t)i * 3) * nqi + a;
        float *gj = t->grad + ((size_t)j * 3) * nqi + a;
        for (int c = 0; c < 3; ++c) {
          const float v = norm * (g * d[c]);
          gi[(size_t)c * nqi] -= v;
          gj[(size_t)c * nqi] += v;
        }
      }
    }
    if (++j == i) {
      ++i;
      j = 0;
    }
  }
  free(sv);
  return NULL;
}

/*
 * fq_out[nqi]: F(Q) at the selected bins (float32 semantics of the
 * reference wrapper).  grad_out (n x 3 x nqi, nullable): -1/2 dF/dq.
 * Returns 0 on success.
 */
int oracle_flat_f32(const float *pos, const float *scat, int64_t n, int nq, float qbin,
                    const int *q_index, int nqi, int nthreads, float *fq_out, float *grad_out) {
  if (n < 2) {
    if (fq_out) memset(fq_out, 0, sizeof(float) * nqi);
    if (grad_out) memset(grad_out, 0, sizeof(float) * n * 3 * nqi);
    return 0;
  }
  if (nthreads < 1) nthreads = 1;
  const int64_t kmax = n * (n - 1) / 2;
  if (nthreads > kmax) nthreads = (int)kmax;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
  task_t *tk = (task_t *)calloc(nthreads, sizeof(task_t));
  const int want_grad = grad_out != NULL;
  for (int t = 0; t < nthreads; ++t) {
    tk[t].pos = pos;
    tk[t].scat = scat;
    tk[t].n = n;
    tk[t].k0 = kmax * t / nthreads;
    tk[t].k1 = kmax * (t + 1) / nthreads;
    tk[t].nq = nq;
    tk[t].nqi = nqi;
    tk[t].q_index = q_index;
    tk[t].qbin = qbin;
    tk[t].want_grad = want_grad;
    tk[t].fsum = (double *)calloc(nqi, sizeof(double));
    tk[t].grad = want_grad ? (float *)calloc((size_t)n * 3 * nqi, sizeof(float)) : NULL;
    pthread_create(&th[t], NULL, worker, &tk[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);

  /* na = N * mean_k(f_i f_j) with the float32 product (closed-form mean in
   * double; the reference takes np.mean of the float32 K x Q array). */
  for (int a = 0; a < nqi; ++a) {
    const int qx = q_index[a];
    double s1 = 0.0, s2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double f = scat[(size_t)i * nq + qx];
      s1 += f;
      s2 += f * f;
    }
    const float na = (float)(((s1 * s1 - s2) / 2.0) / (double)kmax) * (float)n;
    double tot = 0.0;
    for (int t = 0; t < nthreads; ++t) tot += tk[t].fsum[a];
    if (fq_out) {
      float v = (float)tot / na;
      if (!isfinite(v)) v = 0.0f;
      fq_out[a] = 2.0f * v;
    }
    if (want_grad) {
      for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
          float acc = 0.0f;
          for (int t = 0; t < nthreads; ++t) acc += tk[t].grad[((size_t)i * 3 + c) * nqi + a];
          float v = acc / na;
          if (!isfinite(v)) v = 0.0f;
          grad_out[((size_t)i * 3 + c) * nqi + a] = v;
        }
    }
  }
  for (int t = 0; t < nthreads; ++t) {
    free(tk[t].fsum);
    free(tk[t].grad);
  }
  free(tk);
  free(th);
  return 0;
}

/*
 * FP64 sums on a subset of Q bins (and of gradient rows), for checks at sizes
 * where the numpy oracle is too slow.  Same mathematics as oracle/scatter.py
 * fq / grad_fq (the FP64 restatement of cpu_flat.py:77-153):
 *   fsum[a]          = sum_{i>j} f_i(Q_a) f_j(Q_a) sin(Q_a r_ij) / r_ij
 *   gsum[row][c][a]  = sum_{j != i} f_i f_j (Q_a cos(Q_a r) - sin(Q_a r)/r) / r^2 (q_i - q_j)_c
 * with i = rows[row]; the caller applies 2/na and -1/na.  f_e(Q_a) comes from
 * ftab[e][a] (n_elem x nqs) through elem[i].
 */
typedef struct {
  const double *pos, *ftab, *Q;
  const int *elem, *rows;
  int64_t n, k0, k1;
  int nqs, r0, r1;
  double *fsum; /* private, nqs */
  double *gsum; /* shared, rows [r0, r1) owned by this task */
} task64_t;

static void *worker64(void *arg) {
  task64_t *t = (task64_t *)arg;
  const int nqs = t->nqs;
  int64_t i, j;
  if (t->k1 > t->k0) {
    k_to_ij(t->k0, &i, &j);
    for (int64_t k = t->k0; k < t->k1; ++k) {
      const double dx = t->pos[3 * i] - t->pos[3 * j], dy = t->pos[3 * i + 1] - t->pos[3 * j + 1],
                   dz = t->pos[3 * i + 2] - t->pos[3 * j + 2];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      const double *fi = t->ftab + (size_t)t->elem[i] * nqs;
      const double *fj = t->ftab + (size_t)t->elem[j] * nqs;
      for (int a = 0; a < nqs; ++a) t->fsum[a] += fi[a] * fj[a] * sin(t->Q[a] * r) / r;
      if (++j == i) {
        ++i;
        j = 0;
      }
    }
  }
  for (int row = t->r0; row < t->r1; ++row) {
    const int64_t ii = t->rows[row];
    double *g = t->gsum + (size_t)row * 3 * nqs;
    const double *fi = t->ftab + (size_t)t->elem[ii] * nqs;
    for (int64_t jj = 0; jj < t->n; ++jj) {
      if (jj == ii) continue;
      const double d[3] = {t->pos[3 * ii] - t->pos[3 * jj], t->pos[3 * ii + 1] - t->pos[3 * jj + 1],
                           t->pos[3 * ii + 2] - t->pos[3 * jj + 2]};
      const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      const double r = sqrt(r2);
      const double *fj = t->ftab + (size_t)t->elem[jj] * nqs;
      for (int a = 0; a < nqs; ++a) {
        const double qr = t->Q[a] * r;
        const double w = fi[a] * fj[a] * (t->Q[a] * cos(qr) - sin(qr) / r) / r2;
        for (int c = 0; c < 3; ++c) g[(size_t)c * nqs + a] += w * d[c];
      }
    }
  }
  return NULL;
}

int oracle_sums_f64(const double *pos, const int *elem, const double *ftab, const double *Q,
                    int64_t n, int nqs, const int *rows, int nrows, int nthreads, double *fsum,
                    double *gsum) {
  if (n < 0 || nqs <= 0 || nthreads < 1 || nrows < 0) return -1;
  const int64_t kmax = n * (n - 1) / 2;
  task64_t *tasks = (task64_t *)calloc((size_t)nthreads, sizeof(task64_t));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  double *priv = (double *)calloc((size_t)nthreads * nqs, sizeof(double));
  if (!tasks || !th || !priv) return -2;
  memset(fsum, 0, sizeof(double) * nqs);
  if (gsum) memset(gsum, 0, sizeof(double) * (size_t)nrows * 3 * nqs);
  for (int t = 0; t < nthreads; ++t) {
    task64_t *k = &tasks[t];
    k->pos = pos;
    k->ftab = ftab;
    k->Q = Q;
    k->elem = elem;
    k->rows = rows;
    k->n = n;
    k->nqs = nqs;
    k->k0 = kmax * t / nthreads;
    k->k1 = kmax * (t + 1) / nthreads;
    k->r0 = gsum ? nrows * t / nthreads : 0;
    k->r1 = gsum ? nrows * (t + 1) / nthreads : 0;
    k->fsum = priv + (size_t)t * nqs;
    k->gsum = gsum;
    pthread_create(&th[t], NULL, worker64, k);
  }
  for (int t = 0; t < nthreads; ++t) {
    pthread_join(th[t], NULL);
    for (int a = 0; a < nqs; ++a) fsum[a] += priv[(size_t)t * nqs + a];
  }
  free(priv);
  free(th);
  free(tasks);
  return 0;
}
