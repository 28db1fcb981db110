// pllc_host.h -- host helpers shared by pllc_runtime.hip and pllc_abi.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "pllc_kernels.h"

struct pllc_params;

void pllc_set_error(const char* where, const char* what);
bool pllc_hip_ok(const char* where, hipError_t e);
void pllc_fill_dims(const pllc_params* p, pllc::Dims& d, pllc::Phys& ph);
template <typename T>
bool pllc_ctx_tables(pllc::Ctx<T>& c, std::vector<void*>& owned, const char* where);
void pllc_dct_matrix(int n, bool inverse, double* m);
void pllc_eigenvalues(int n, double h, double* lam);
void pllc_set_nmvp(pllc::Phys& ph, double M, double Mum, double cm);
double pllc_slope_parameter(const double* c1d, int n, double M, double Mum, double Lambda);
void pllc_centre_line(const double* Yc, int js, int je, const double* Zc, int ks, int ke, int* jj, int* kk);

// the learned closure (pllc_closure.hip)
struct pllc_cnn;
template <typename T>
void pllc_cnn_correct(pllc_cnn* m, const pllc::Dims& d, const T* c, const T* u, const T* v, const T* w, const T* mu,
                      T* ncorr, hipStream_t st);
bool pllc_cnn_dims_match(const pllc_cnn* m, const pllc::Dims& d);
template <typename T>
bool pllc_guard_stats(const pllc::Dims& d, const T* c, const T* ncorr, double* scratch, double* out, hipStream_t st,
                      const char* W);
