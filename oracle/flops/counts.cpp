// counts.cpp -- counter storage of the FLOP-counting oracle build
// (fdouble.hpp).  TEST INFRASTRUCTURE ONLY.
extern "C" {
unsigned long long oracle_flop_counts[2][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
int oracle_flop_share = 0;

// out[10]: additions, subtractions, multiplications, divisions, libm calls
// outside the finite-difference functions, then the same inside them
void oracle_flops_get(unsigned long long* out) {
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < 5; ++i) out[5 * s + i] = oracle_flop_counts[s][i];
}
void oracle_flops_reset(void) {
  for (int s = 0; s < 2; ++s)
    for (int i = 0; i < 5; ++i) oracle_flop_counts[s][i] = 0;
  oracle_flop_share = 0;
}
}
