/* ref_soft_shim.c -- thin driver around soft-1.0 routines compiled from the reference sources into
 * oracle/_ref/libsoftref.so.  TEST INFRASTRUCTURE ONLY: it checks the oracle's S^2/SO(3) restatement.
 * It follows the call sequence of nbfProjectionRotationMetric3D::soft (ProjRot.h:1137-1295) and ::harmonics
 * (:1457-1554), using the non-FFTW variants so3CombineCoef (so3_correlate_sym.c:100) and
 * Inverse_SO3_Naive_sym (soft_sym.c:1162), which compute the same transform as the FFTW ones ETTK calls.
 */
#include <stdlib.h>
#include <string.h>
#include "complex.h" /* soft-1.0 header: defines REAL */
#include "FST_semi_memo.h"
#include "cospmls.h"
#include "primitive_FST.h"
#include "so3_correlate_sym.h"
#include "soft_sym.h"
#include "rotate_so3.h"
#include "utils_so3.h"

static double **make_table(int bw, double **space) {
    *space = (double *)malloc(sizeof(double) * (Reduced_Naive_TableSize(bw, bw) + Reduced_SpharmonicTableSize(bw, bw)));
    double *ws = (double *)malloc(sizeof(double) * (16 * bw * bw + 64 * bw));
    double **t = SemiNaive_Naive_Pml_Table(bw, bw, *space, ws);
    free(ws);
    return t;
}

/* S^2 forward transform of real samples f (2bw x 2bw, [theta][phi]); coefficients in SOFT order */
void ref_fst(int bw, const double *f, double *cr, double *ci) {
    int n = 2 * bw;
    double *space, **tab = make_table(bw, &space);
    double *fr = (double *)malloc(sizeof(double) * n * n), *fi = (double *)calloc(n * n, sizeof(double));
    double *ws = (double *)malloc(sizeof(double) * (16 * bw * bw + 64 * bw));
    memcpy(fr, f, sizeof(double) * n * n);
    FST_semi_memo(fr, fi, cr, ci, n, tab, ws, 1, bw);
    free(ws); free(fi); free(fr); free(tab); free(space);
}

int ref_seanindex(int m, int l, int bw) { return seanindex(m, l, bw); }

/* correlation output in SOFT's [beta][alpha][gamma] layout, real part */
void ref_soft_correlate(int bw, const double *sig, const double *pat, double *out) {
    int n = 2 * bw;
    double *space, **tab = make_table(bw, &space);
    double *fr = (double *)malloc(sizeof(double) * n * n), *fi = (double *)calloc(n * n, sizeof(double));
    double *sR = malloc(sizeof(double) * bw * bw), *sI = malloc(sizeof(double) * bw * bw);
    double *pR = malloc(sizeof(double) * bw * bw), *pI = malloc(sizeof(double) * bw * bw);
    double *ws2 = (double *)malloc(sizeof(double) * (16 * bw * bw + 64 * bw));
    int nc = totalCoeffs_so3(bw);
    double *cR = malloc(sizeof(double) * nc), *cI = malloc(sizeof(double) * nc);
    double *oR = malloc(sizeof(double) * n * n * n), *oI = malloc(sizeof(double) * n * n * n);
    double *ws1 = malloc(sizeof(double) * 16 * bw * bw * bw);
    memcpy(fr, sig, sizeof(double) * n * n);
    FST_semi_memo(fr, fi, sR, sI, n, tab, ws2, 1, bw);
    memcpy(fr, pat, sizeof(double) * n * n);
    memset(fi, 0, sizeof(double) * n * n);
    FST_semi_memo(fr, fi, pR, pI, n, tab, ws2, 1, bw);
    so3CombineCoef(bw, bw, bw - 1, sR, sI, pR, pI, cR, cI);
    Inverse_SO3_Naive_sym(bw, cR, cI, oR, oI, ws1, ws2, 1);
    memcpy(out, oR, sizeof(double) * n * n * n);
    free(ws1); free(oI); free(oR); free(cI); free(cR); free(ws2);
    free(pI); free(pR); free(sI); free(sR); free(fi); free(fr); free(tab); free(space);
}

/* inverse SO(3) transform of given coefficients (SOFT order), real and imaginary output [beta][alpha][gamma] */
void ref_inverse_so3(int bw, const double *cR, const double *cI, double *oR, double *oI) {
    double *ws1 = malloc(sizeof(double) * 16 * bw * bw * bw);
    double *ws2 = malloc(sizeof(double) * (16 * bw * bw + 64 * bw));
    double *r = malloc(sizeof(double) * totalCoeffs_so3(bw)), *i = malloc(sizeof(double) * totalCoeffs_so3(bw));
    memcpy(r, cR, sizeof(double) * totalCoeffs_so3(bw));
    memcpy(i, cI, sizeof(double) * totalCoeffs_so3(bw));
    Inverse_SO3_Naive_sym(bw, r, i, oR, oI, ws1, ws2, 1);
    free(i); free(r); free(ws2); free(ws1);
}

int ref_so3_coef_loc(int m1, int m2, int l, int bw) { return so3CoefLoc(m1, m2, l, bw); }
int ref_total_coeffs_so3(int bw) { return totalCoeffs_so3(bw); }

/* rotateFct(bw, bw, bw-1, ...) as nbfProjectionRotationMetric3D::harmonics calls it */
void ref_rotate(int bw, const double *in, double *out, double alpha, double beta, double gamma) {
    int n = 2 * bw;
    double *s1, **t1 = make_table(bw, &s1);
    double *s2, **t2 = make_table(bw, &s2);
    double *s3 = (double *)malloc(sizeof(double) * (Reduced_Naive_TableSize(bw, bw) + Reduced_SpharmonicTableSize(bw, bw)));
    double *scratch = (double *)malloc(sizeof(double) * (16 * bw * bw + 64 * bw));
    double **tt = Transpose_SemiNaive_Naive_Pml_Table(t2, bw, bw, s3, scratch);
    double *iR = malloc(sizeof(double) * n * n), *iI = calloc(n * n, sizeof(double));
    double *oI = malloc(sizeof(double) * n * n);
    memcpy(iR, in, sizeof(double) * n * n);
    rotateFct(bw, bw, bw - 1, iR, iI, out, oI, alpha, beta, gamma, scratch, t1, tt);
    free(oI); free(iI); free(iR); free(tt); free(scratch); free(s3); free(t2); free(s2); free(t1); free(s1);
}
