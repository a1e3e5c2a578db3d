/* fem_oracle.h -- CPU restatement of the reference algorithm (test infrastructure only). */
#ifndef FEM_ORACLE_H
#define FEM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int64_t orc_pattern(int npe, int64_t nv, int64_t ne, const int32_t* conn, int64_t* rp, int32_t* ci);
void orc_gauss_lobatto_jacobi(int q, double a, double b, double* z, double* w);
void orc_gauss_radau_jacobi(int q, double a, double b, double* z, double* w);
void orc_tet_mass_integrals(double out[10]);
void orc_tri_mass_integrals(double out[6]);
void orc_tet_element(const double* X, double co, const double integ[10], double K[16], double M[16]);
void orc_tri_element(const double* X, int signed_mass, const double integ[6], double K[9], double M[9]);
int orc_assemble(int dim, int64_t nv, const double* xyz, int64_t ne, const int32_t* conn, const double* kappa,
                 double mass_coeff, int signed_mass, const int64_t* rp, const int32_t* ci, double* vals);
void orc_element_grad(int dim, const double* xyz, int64_t ne, const int32_t* conn, const double* a, const double* u,
                      double* out);
#ifdef __cplusplus
}
#endif
#endif
#ifdef __cplusplus
extern "C" {
#endif
uint32_t orc_mis_seed(uint64_t seed, uint32_t i);
void orc_mis(int64_t n, const int32_t* ap, const int32_t* aj, int depth, uint64_t seed, int32_t* mis);
int64_t orc_aggregate(int64_t n, const int32_t* ap, const int32_t* aj, int depth, int min_size, int rule,
                      uint64_t seed, int32_t* agg);
#ifdef __cplusplus
}
#endif
#ifdef __cplusplus
extern "C" {
#endif
void orc_induced_graph(int64_t n, const int32_t* gp, const int32_t* gj, const int32_t* agg, int64_t nc, int32_t** cp,
                       int32_t** cj);
void orc_free(void* p);
int orc_amg_pcg(int64_t n, const int64_t* rp, const int32_t* ci, const double* val, const double* b, double* x,
                double tol, int max_iters, int top_size, int max_levels, int pre, int post, int relaxes,
                double smoother_weight, double pro_omega, int part_max, uint64_t seed, double* out);
#ifdef __cplusplus
}
#endif
#ifdef __cplusplus
extern "C" {
#endif
int64_t orc_condition(int64_t n, const int32_t* ap, const int32_t* aj, int32_t* agg, int64_t nagg, int min_size,
                      int max_size);
/* aggregatorType 4 / 5 (fem_oracle_aggmis.c): flood-fill distance-k MIS from `starter`
 * (AggMIS_MIS_CPU.cu:25-140) and aggregation to the nearest root (AggMIS_Aggregation_CPU.cu:4-92) */
void orc_flood_fill_mis(int64_t n, const int32_t* ap, const int32_t* aj, int k, int64_t starter, int32_t* mis);
int64_t orc_aggregate_to_nearest(int64_t n, const int32_t* ap, const int32_t* aj, const int32_t* roots,
                                 int32_t* agg);
#ifdef __cplusplus
}
#endif
