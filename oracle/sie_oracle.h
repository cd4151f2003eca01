// ORACLE — test infrastructure only (see sie_oracle.cpp). C ABI of the CPU
// FP64 restatement used by tests/ and the bench's CPU baseline.
#pragma once
#ifdef __cplusplus
extern "C" {
#endif
int oracle_rule_order(int rule);
void oracle_rule(int rule, float* xy, float* w);
void oracle_pair_regular(const double* tt, const int* ti, const double* st, const int* si, int rule,
                         double k1, double k2, int singular, double* out);
void oracle_pair_singular(const double* tt, const int* ti, const double* st, const int* si, int rule,
                          double k1, double k2, double* out);
void oracle_wilton(const double* r, const double* tri, double* out);
void oracle_mass_batch(const double* tri, const int* idx, int rule, const double* T, double* out);
void oracle_excitation_planewave(const double* tri, const int* idx, int rule, const double* E0,
                                 const double* kh, double k, double eta, double* VE, double* VH);
void oracle_face_basis(const double* tri, int basis, double* u, double* v);
int oracle_icosphere(float radius, int level, double* V, int* F, int* E, int* FE, int* counts);
void oracle_pair_blocks_all(const double* V, const int* F, int nf, int rule, float k1, float k2,
                            float thr, double* out);
void oracle_assemble_gstc(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                          float k1, float k2, float mu_r1, float mu_r2, float eps_r1, float eps_r2,
                          float mu0, float eps0, float omega, float thr, const double* tensors,
                          int uniform, int basis, int with_calderon, int with_mass, double* Z);
void oracle_assemble_gstc_ex(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                             float k1, float k2, float mu_r1, float mu_r2, float eps_r1, float eps_r2,
                             float mu0, float eps0, float omega, float thr, const double* tensors,
                             int uniform, int basis, int with_calderon, int with_mass, int scaled, double* Z);
void oracle_gstc_rows(const double* V, const int* F, const int* FE, int nf, int ne, int rule, float k1, float k2,
                      float mu_r1, float mu_r2, float eps_r1, float eps_r2, float mu0, float eps0, float omega,
                      float thr, const double* tensors, int uniform, int basis, int nrows, const int* rows,
                      double* out);
void oracle_fields(const double* V, const int* F, const int* FE, int nf, int rule, const double* J,
                   const double* M, const double* obs, int nobs, double omega, double eps, double mu, double k,
                   double* E, double* H);
void oracle_assemble_pec(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                         float k2, float mu_r2, float mu0, float omega, float thr, double* Z);
void oracle_assemble_pmchwt(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                            float k1, float k2, float mu_r1, float mu_r2, float eps_r1, float eps_r2,
                            float mu0, float eps0, float omega, float thr, double* Z);
void oracle_rhs_planewave(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                          const double* E0, const double* kh, double k, double eta, int region,
                          int kind, float mu0, float eps0, double vac_mu, double vac_eps, double* rhs);
#ifdef __cplusplus
}
#endif
