// Drop-in replacement for the reference's src/linsys/HelmholtzEqn.h: the same
// public interface (constructor, solField/rhsField/c/nu/dt/svvCutoff,
// Assemble, Solve) without the PETSc members, solved by libse_amd on the GPU.
// A = c M + nu dt K~ (+ dt SVV with svvCutoff > 0) with the reference's element
// operators (src/linsys/HelmholtzEqn.cpp:21-82); Solve assembles f = M rhs + b,
// solves A x = f by Jacobi-PCG to a true relative residual of 1e-10 (SE_RTOL)
// and writes x back into solField (HelmholtzEqn.cpp:105-136).  Fields of 1 or
// 2 dofs (a 2-dof field solves one operator per dof, with that dof's Dirichlet
// sides).  SE_PRECISION=fp32 or bf16 selects the mixed-precision solver
// (operators without SVV); fp64_fp32c / fp64_bf16c the FDM-preconditioned CG
// with an FP64 first application and FP32 / BF16 corrections.  Include after
// the reference's Field.h, like the original.
#ifndef SE_DROPIN_HELMHOLTZ_EQN_H
#define SE_DROPIN_HELMHOLTZ_EQN_H
#include <string>

#include "se_amd.h"

class Field;

class HelmholtzEqn {
   public:
    HelmholtzEqn(Field* _solField, Field* _rhsField, double _c, double _nu, double _dt, int _svvCutoff);
    ~HelmholtzEqn();
    Field* solField;
    Field* rhsField;
    double c;
    double nu;
    double dt;
    int svvCutoff;
    void Assemble();
    void Solve(std::string name, bool removeNullSpace = false);
    se_solve_info info;  // outcome of the last Solve

   private:
    bool assembled;
    se_op* op;   // dof 0
    se_op* op2;  // dof 1 of a 2-dof field
};

#endif
