/* C-ABI of libPriorFunctionSolver.so (MI355X build): the TV-SQS part of the
 * reference's PriorFunctionSolver library.  Host pointers, synchronous,
 * layout [nb][nx][ny][(nz)][nc]; never throws (failures print
 * "<function> failed: <reason>" and set pfs_last_error()). */
#ifndef PRIORFUNCTIONSOLVER_H
#define PRIORFUNCTIONSOLVER_H

#ifdef __cplusplus
extern "C" {
#endif

/* device selector of the reference library (python/PriorFunctionSolver.py:73-74) */
int cSetDevice(int device);
/* PriorFunctionSolver/TVSQS.cu:108 */
void cTVSQS2D(float* s1, float* s2, float* var, const float* img, int nb, int nx, int ny, int nc, float eps);
/* PriorFunctionSolver/TVSQS.cu:290 */
void cTVSQS3D(float* s1, float* s2, float* var, const float* img, int nb, int nx, int ny, int nz, int nc, float eps);
/* additive: message of the last failure on the calling thread ("" after success) */
const char* pfs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
