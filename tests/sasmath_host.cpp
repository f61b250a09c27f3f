// Host build of sasmodels_amd/csrc/sasmath.h for tests/test_sasmath.py: the
// FP32 special functions of the kernels, compiled for the CPU.
#include "../sasmodels_amd/csrc/sasmath.h"

extern "C" {
double t_sin_f(double x) { return sas::sin_f((float)x); }
double t_sinc_f(double x) { return sas::sinc_f((float)x); }
double t_sph_j1c_f(double x) { return sas::sph_j1c_f((float)x); }
double t_j1c_f(double x) { return sas::j1c_f((float)x); }
double t_sinc_d(double x) { return sas::sinc_d(x); }
double t_sph_j1c_d(double x) { return sas::sph_j1c_d(x); }
double t_j1c_d(double x) { return sas::j1c_d(x); }
double t_sph_j1c_dd(double x) { return sas::sph_j1c_dd(x); }
}
