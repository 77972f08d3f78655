// Drop-in replacement for the reference's src/linsys/PoissonEqn.h: A = c K~
// (src/linsys/PoissonEqn.cpp:30-80), Solve as HelmholtzEqn's
// (PoissonEqn.cpp:102-134), removeNullSpace projecting out the constant
// (periodic Poisson).  Include after the reference's Field.h.
#ifndef SE_DROPIN_POISSON_EQN_H
#define SE_DROPIN_POISSON_EQN_H
#include "se_amd.h"

class Field;

class PoissonEqn {
   public:
    PoissonEqn(Field* _solField, Field* _rhsField, double _c);
    ~PoissonEqn();
    Field* solField;
    Field* rhsField;
    double c;
    void Assemble();
    void Solve(bool removeNullSpace);
    se_solve_info info;  // outcome of the last Solve

   private:
    bool assembled;
    se_op* op;
};

#endif
