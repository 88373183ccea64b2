// fr_host.h -- host-side geometry shared by the C-ABI glue.
#pragma once

namespace fr {
float lcg_random(int* is);
void matmul_t(const float a[9], const float b[9], float c[9]);
void rotation(float phi, float theta, float psi, float amag, const float* sm, float dm[9]);
int point_group(char asym, int nasym, float* symop, int* jsym, int* nsym, int* iredun);
int symmetry_copies(const float tm[9], int nsym, const int* jsym, const float* symop, float* out);
int search_grid(float rmm, float ri, float dangin, char asym, int nasym, const int mask[5],
                int asymstore, float* sang, int cap, int* iquadmax);
}  // namespace fr
