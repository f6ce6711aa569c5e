/*
 * CPU restatement of the reference's per-row three-phonon loop, used as the timed CPU
 * baseline and as a checker.  TEST INFRASTRUCTURE ONLY: nothing in the product links it.
 *
 * Follows src/thermal_conductivity/scattering_threephonon.f90:
 *   pretransform_phi3 (:275-317) recomputed for every (row, triplet), as the reference does;
 *   evp1 = zgeru(egv2, egv1) (:97); evp2 = conj(zgeru(egv3, evp1)) (:109-110);
 *   c0 = dot_product(evp2, ptf) (:112); psisq (:113); get_dirac (:219-271) for the
 *   adaptive Gaussian; f0 + f1 accumulated per equivalent triplet (:116-128).
 * Only the linewidth is accumulated (the off-diagonal bookkeeping is not timed).
 * Rows are spread over OpenMP threads like the reference spreads them over MPI ranks.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double gaussf(double x, double mu, double sigma) {
    double d = x - mu;
    if (fabs(d) > 8.0 * sigma) return 0.0;
    double inv = 1.0 / sigma;
    return exp(-d * d * 0.5 * inv * inv) * 0.398942280401433 * inv;
}

/*
 * nb bands, nq full points, ntf third-order IFCs.
 * fc3: a1, a2, a3 (ntf), lv2, lv3 (3*ntf), mwm (27*ntf, [i][j][k]).
 * qcart (3*nq), om_full (nq*nb), egv_full (nq*nb*nb complex, [q][component][mode]),
 * irr_of_full (nq), be (nirr*nb), sigsq (nirr*nb).
 * Rows: nrow entries of (q1 irreducible, b1, om1, egv1 column (nb complex, component order)).
 * For each row its triplet list: trip_off (nrow+1), tq2, tq3, perm, neq.
 * Output gamma (nrow).  Returns the number of (b2, b3) elements formed.
 */
long ref_rows(int nb, int nq, int ntf, const int *a1, const int *a2, const int *a3, const double *lv2,
              const double *lv3, const double *mwm, const double *qcart, const double *om_full,
              const double _Complex *egv_full, const int *irr_of_full, const double *be, const double *sigsq,
              int nrow, const int *row_q1, const int *row_b1, const double *row_om1,
              const double _Complex *row_egv1, const int *trip_off, const int *tq2, const int *tq3,
              const double *perm, const int *neq, double mcweight, double *gamma) {
    const double pi = 3.141592653589793, twopi = 6.283185307179586, freqtol = 1e-9;
    long total = 0;
#pragma omp parallel reduction(+ : total)
    {
        const long nb2 = (long)nb * nb, nb3 = nb2 * nb;
        double _Complex *ptf = malloc(sizeof(double _Complex) * nb3);
        double _Complex *evp1 = malloc(sizeof(double _Complex) * nb2);
        double _Complex *evp2 = malloc(sizeof(double _Complex) * nb3);
        double _Complex *e1 = malloc(sizeof(double _Complex) * nb);
        double _Complex *e2 = malloc(sizeof(double _Complex) * nb);
        double _Complex *e3 = malloc(sizeof(double _Complex) * nb);
#pragma omp for schedule(dynamic, 1)
        for (int r = 0; r < nrow; ++r) {
            const int q1 = row_q1[r], b1 = row_b1[r];
            const double om1 = row_om1[r];
            double g0 = 0.0;
            for (int i = 0; i < nb; ++i) e1[i] = row_egv1[(long)r * nb + i] / sqrt(om1);
            for (int it = trip_off[r]; it < trip_off[r + 1]; ++it) {
                const int q2 = tq2[it], q3 = tq3[it];
                const double *Q2 = qcart + 3 * q2, *Q3 = qcart + 3 * q3;
                memset(ptf, 0, sizeof(double _Complex) * nb3);
                for (int t = 0; t < ntf; ++t) {
                    double iqr = Q2[0] * lv2[3 * t] + Q2[1] * lv2[3 * t + 1] + Q2[2] * lv2[3 * t + 2] +
                                 Q3[0] * lv3[3 * t] + Q3[1] * lv3[3 * t + 1] + Q3[2] * lv3[3 * t + 2];
                    iqr = -iqr * twopi;
                    const double _Complex ex = cos(iqr) + I * sin(iqr);
                    for (int i = 0; i < 3; ++i)
                        for (int j = 0; j < 3; ++j)
                            for (int k = 0; k < 3; ++k) {
                                const long ia = 3 * a1[t] + i, ib = 3 * a2[t] + j, ic = 3 * a3[t] + k;
                                ptf[ia * nb2 + ib * nb + ic] += mwm[27 * t + 9 * i + 3 * j + k] * ex;
                            }
                }
                const int i2 = irr_of_full[q2], i3 = irr_of_full[q3];
                for (int b2 = 0; b2 < nb; ++b2) {
                    const double om2 = om_full[(long)q2 * nb + b2];
                    if (om2 < freqtol) continue;
                    if (fabs(om1 - om2) < freqtol) continue;
                    for (int i = 0; i < nb; ++i) e2[i] = egv_full[((long)q2 * nb + i) * nb + b2] / sqrt(om2);
                    for (int j = 0; j < nb; ++j)
                        for (int i = 0; i < nb; ++i) evp1[(long)j * nb + i] = e2[i] * e1[j];
                    for (int b3 = 0; b3 < nb; ++b3) {
                        const double om3 = om_full[(long)q3 * nb + b3];
                        if (om3 < freqtol) continue;
                        if (fabs(om1 - om3) < freqtol || fabs(om2 - om3) < freqtol) continue;
                        for (int i = 0; i < nb; ++i) e3[i] = egv_full[((long)q3 * nb + i) * nb + b3] / sqrt(om3);
                        const double sigma = sqrt(sigsq[(long)q1 * nb + b1] + sigsq[(long)i2 * nb + b2] +
                                                  sigsq[(long)i3 * nb + b3]);
                        const double d0 = gaussf(om1, -om2 + om3, sigma) - gaussf(om1, om2 - om3, sigma);
                        const double d1 = gaussf(om1, om2 + om3, sigma) - gaussf(om1, -om2 - om3, sigma);
                        for (long j = 0; j < nb2; ++j)
                            for (int i = 0; i < nb; ++i) evp2[j * nb + i] = e3[i] * evp1[j];
                        for (long l = 0; l < nb3; ++l) evp2[l] = conj(evp2[l]);
                        double _Complex c0 = 0;
                        for (long l = 0; l < nb3; ++l) c0 += conj(evp2[l]) * ptf[l];
                        const double psisq = pi / 16.0 * creal(c0 * conj(c0)) * mcweight;
                        const double n2 = be[(long)i2 * nb + b2], n3 = be[(long)i3 * nb + b3];
                        const double f0 = perm[it] * psisq * (n2 - n3) * d0;
                        const double f1 = perm[it] * psisq * (n2 + n3 + 1.0) * d1;
                        for (int e = 0; e < neq[it]; ++e) g0 = g0 + f0 + f1;
                        ++total;
                    }
                }
            }
            gamma[r] = g0;
        }
        free(ptf); free(evp1); free(evp2); free(e1); free(e2); free(e3);
    }
    return total;
}

int ref_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
