// Drop-in replacement for the reference's src/base/Advector.h: first-order
// (two-argument constructor) and second-order (four-argument constructor)
// semi-Lagrangian advection of `field` by `velocity` over dt into `fieldSL`
// (src/base/Advector.cpp:26-349), run by libse_amd on the GPU; Interpolate
// is the bicubic point interpolation (Advector.cpp:278-337), also on the GPU;
// cubicInterp selects it for the advected field (default N < 3).  The
// second-order step leaves the extrapolated velocities in velPlusOne,
// velPlusHalf, velMinusHalf as the reference does.  Include after the
// reference's Field.h.
#ifndef SE_DROPIN_ADVECTOR_H
#define SE_DROPIN_ADVECTOR_H

class Field;

class Advector {
   public:
    Advector(Field* _field, Field* _velocity);
    Advector(Field* _field, Field* _velocity, Field* _fieldMinusOne, Field* _velMinusOne);
    ~Advector();
    void Advect(double dt);
    void Interpolate(Field* field, double* coord, double* psi);
    Field* field;
    Field* velocity;
    Field* fieldMinusOne;
    Field* velMinusOne;
    Field* velPlusOne;
    Field* velPlusHalf;
    Field* velMinusHalf;
    Field* fieldSL;
    bool secondOrder;
    bool cubicInterp;
};

#endif
