// Test probe: the product's portable math and MT19937 / uniform draws
// (gpu-smina_amd/csrc/pmath.h, rng.h) exposed to the CPU tests.
#include <stdint.h>
#include "../../gpu-smina_amd/csrc/pmath.h"
#include "../../gpu-smina_amd/csrc/rng.h"

extern "C" {
void pm_sin(const double* x, double* y, int n) { for (int i = 0; i < n; i++) y[i] = pmath::sin(x[i]); }
void pm_cos(const double* x, double* y, int n) { for (int i = 0; i < n; i++) y[i] = pmath::cos(x[i]); }
void pm_exp(const double* x, double* y, int n) { for (int i = 0; i < n; i++) y[i] = pmath::exp(x[i]); }
void pm_acos(const double* x, double* y, int n) { for (int i = 0; i < n; i++) y[i] = pmath::acos(x[i]); }
void mt_stream(uint32_t seed, uint32_t* out, int n)
{
	sgpu::MTState s;
	sgpu::mt_seed(s, seed);
	for (int i = 0; i < n; i++)
		out[i] = sgpu::mt_next(s);
}
void uniform_fl(uint32_t seed, double a, double b, double* out, int n)
{
	sgpu::MTState s;
	sgpu::mt_seed(s, seed);
	sgpu::MTRef g{&s};
	for (int i = 0; i < n; i++)
		out[i] = sgpu::random_fl(a, b, g);
}
void uniform_int(uint32_t seed, int a, int b, int* out, int n)
{
	sgpu::MTState s;
	sgpu::mt_seed(s, seed);
	sgpu::MTRef g{&s};
	for (int i = 0; i < n; i++)
		out[i] = sgpu::random_int(a, b, g);
}
}
