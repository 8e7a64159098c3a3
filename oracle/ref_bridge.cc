// TEST INFRASTRUCTURE ONLY.  Exposes the reference's own Newton solver
// (dengo/solvers/BE_chem_solve.C:79-83, compiled unmodified from the reference
// tree by oracle/Makefile into oracle/_ref/) under a C name, so that tests can
// hand it to the oracle's substep driver in place of the oracle's restatement.
typedef int (*rhs_f)(double *, double *, int, int, void *);
typedef int (*jac_f)(double *, double *, int, int, void *);
int BE_chem_solve(rhs_f f, jac_f J, double *u, double dt, double *rtol, double *atol,
                  int nstrip, int nchem, double *scaling, void *sdata, double *u0,
                  double *s, double *gu, double *Ju);

extern "C" int ref_BE_chem_solve(rhs_f f, jac_f J, double *u, double dt, double *rtol,
                                 double *atol, int nstrip, int nchem, double *scaling,
                                 void *sdata, double *u0, double *s, double *gu,
                                 double *Ju)
{
    return BE_chem_solve(f, J, u, dt, rtol, atol, nstrip, nchem, scaling, sdata, u0, s,
                         gu, Ju);
}
