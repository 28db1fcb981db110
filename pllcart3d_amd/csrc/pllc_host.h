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
