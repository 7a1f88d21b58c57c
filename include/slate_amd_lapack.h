/*
 * slate_amd_lapack.h — LAPACK-compatible entry points of libslate_amd.so.
 *
 * Fortran calling convention (all arguments by pointer), HOST column-major
 * arrays, 32-bit integers, 1-based ipiv, LAPACK argument order and info
 * semantics (info = -i for an illegal i-th argument; info > 0 singular /
 * not positive definite; info <= -1000 an internal error).  They stage the
 * arrays through device memory around one call of include/slate_amd.h.
 * Tile sizes come from SLATE_LAPACK_NB and SLATE_LAPACK_IB when set
 * (lapack_api/lapack_slate.hh).
 */
#ifndef SLATE_AMD_LAPACK_H
#define SLATE_AMD_LAPACK_H

#ifdef __cplusplus
extern "C" {
#endif

/* lapack_api/lapack_getrf.cc (slate_dgetrf / slate_sgetrf); square only. */
void slate_dgetrf_(const int* m, const int* n, double* a, const int* lda, int* ipiv, int* info);
void slate_sgetrf_(const int* m, const int* n, float* a, const int* lda, int* ipiv, int* info);

/* lapack_api/lapack_getrs.cc (slate_dgetrs / slate_sgetrs). */
void slate_dgetrs_(const char* trans, const int* n, const int* nrhs, const double* a, const int* lda,
                   const int* ipiv, double* b, const int* ldb, int* info);
void slate_sgetrs_(const char* trans, const int* n, const int* nrhs, const float* a, const int* lda,
                   const int* ipiv, float* b, const int* ldb, int* info);

/* lapack_api/lapack_gesv.cc:32-35 (slate_dgesv). */
void slate_dgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
                  const int* ldb, int* info);

/* lapack_api/lapack_gesv_mixed.cc:24 (slate_dsgesv); work/swork unused. */
void slate_dsgesv_(const int* n, const int* nrhs, double* a, const int* lda, int* ipiv, double* b,
                   const int* ldb, double* x, const int* ldx, double* work, float* swork, int* iter,
                   int* info);

/* lapack_api/lapack_potrf.cc (slate_dpotrf / slate_spotrf). */
void slate_dpotrf_(const char* uplo, const int* n, double* a, const int* lda, int* info);
void slate_spotrf_(const char* uplo, const int* n, float* a, const int* lda, int* info);

/* LAPACK dpotrs argument list over slate::potrs (src/potrs.cc:55-77). */
void slate_dpotrs_(const char* uplo, const int* n, const int* nrhs, const double* a, const int* lda,
                   double* b, const int* ldb, int* info);

/* lapack_api/lapack_posv.cc (slate_dposv). */
void slate_dposv_(const char* uplo, const int* n, const int* nrhs, double* a, const int* lda, double* b,
                  const int* ldb, int* info);

/* LAPACK dsposv argument list over slate::posv_mixed (src/posv_mixed.cc:105-292). */
void slate_dsposv_(const char* uplo, const int* n, const int* nrhs, double* a, const int* lda, double* b,
                   const int* ldb, double* x, const int* ldx, double* work, float* swork, int* iter,
                   int* info);

#ifdef __cplusplus
}
#endif

#endif /* SLATE_AMD_LAPACK_H */
