// krylov.hip -- Solve_PBiCGStab (vecxitrsol.h:350-437) on device vectors:
// the reference's recurrences, operation for operation, over the library's
// MatVec, ILU SolvePrecond and vector kernels; each dot product is a host
// read (the method needs five scalars per iteration and the stop test).
// The matrices it serves (stiffness-warped backward Euler) are not
// symmetric, so the fused PCG does not apply.
#include "handles.hpp"

namespace dfm2h {

static void ck(int st) {
  if (st != DFM2H_OK) throw Error(st, dfm2h_last_error());
}

static double dot(const double *x, const double *y, size_t n) {
  double d = 0.0;
  ck(dfm2h_vec_dot(x, y, n, &d));
  return d;
}

}  // namespace dfm2h

using namespace dfm2h;

extern "C" {

int dfm2h_solve_pbicgstab(dfm2h_bsr A, dfm2h_ilu P, double *r, double *x, double tol,
                          unsigned max_niter, double *hist, size_t hist_cap, size_t *n_hist) {
  return guard(__func__, [&] {
    DFM2H_REQUIRE(A && P && r && x && n_hist, DFM2H_ERR_INVALID, "null argument");
    DFM2H_REQUIRE(hist || hist_cap == 0, DFM2H_ERR_INVALID, "null history with a capacity");
    size_t nblk = 0;
    ck(dfm2h_bsr_info(A, &nblk, nullptr, nullptr, nullptr));
    const size_t n = nblk * 3;
    std::vector<double> h;
    auto finish = [&] {
      for (size_t k = 0; k < h.size() && k < hist_cap; ++k) hist[k] = h[k];
      *n_hist = h.size();
    };
    ck(dfm2h_vec_set_zero(x, n));  // {u} = 0
    const double sq_ini = dot(r, r, n);
    if (sq_ini < 1.0e-60) {
      h.push_back(std::sqrt(sq_ini));
      finish();
      return;
    }
    const double sq_inv_ini = 1.0 / sq_ini;
    DBuf<double> s, Ms, AMs, Mp, AMp, r0, p;
    for (DBuf<double> *v : {&s, &Ms, &AMs, &Mp, &AMp, &r0, &p}) {
      v->alloc(n);
      v->zero();
    }
    ck(dfm2h_vec_copy(r0.p, r, n));  // {r2} = {r}
    ck(dfm2h_vec_copy(p.p, r, n));   // {p} = {r}
    for (unsigned it = 1; it < max_niter; ++it) {
      ck(dfm2h_vec_copy(Mp.p, p.p, n));
      ck(dfm2h_ilu_solve_precond(P, Mp.p));
      const double r_r2 = dot(r, r0.p, n);
      ck(dfm2h_bsr_matvec(A, AMp.p, 1.0, Mp.p, 0.0));
      const double alpha = r_r2 / dot(AMp.p, r0.p, n);
      ck(dfm2h_vec_copy(s.p, r, n));
      ck(dfm2h_vec_add_scaled(s.p, -alpha, AMp.p, n));
      ck(dfm2h_vec_copy(Ms.p, s.p, n));
      ck(dfm2h_ilu_solve_precond(P, Ms.p));
      ck(dfm2h_bsr_matvec(A, AMs.p, 1.0, Ms.p, 0.0));
      double omega;
      {
        const double denominator = dot(AMs.p, AMs.p, n);
        const double numerator = dot(s.p, AMs.p, n);
        omega = numerator / denominator;
      }
      ck(dfm2h_vec_add_scaled(x, alpha, Mp.p, n));
      ck(dfm2h_vec_add_scaled(x, omega, Ms.p, n));
      ck(dfm2h_vec_copy(r, s.p, n));
      ck(dfm2h_vec_add_scaled(r, -omega, AMs.p, n));
      const double conv = std::sqrt(dot(r, r, n) * sq_inv_ini);
      h.push_back(conv);
      if (conv < tol) break;
      const double beta = dot(r, r0.p, n) * alpha / (r_r2 * omega);
      // p *= beta; p += 1.0 r  (one kernel: beta p + r, the same two roundings)
      ck(dfm2h_vec_scale_and_add(p.p, beta, r, n));
      ck(dfm2h_vec_add_scaled(p.p, -beta * omega, AMp.p, n));
    }
    finish();
  });
}

}  // extern "C"
