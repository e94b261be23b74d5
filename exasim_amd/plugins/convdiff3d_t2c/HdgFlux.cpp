#include "pdemodel.h"
#include <cmath>

// f = kappa q + c u, f[c + 2 d]; param = (kappa, c_x, c_y, c_z); udg = (u0, u1, q: u_c + 2 d at 2 + c + 2 d)
void HdgFlux(dstype* f, dstype* f_udg, dstype* f_wdg, const dstype* xdg, const dstype* udg, const dstype* odg,
             const dstype* wdg, const dstype* uinf, const dstype* param, const dstype time, const int modelnumber,
             const int ng, const int nc, const int ncu, const int nd, const int ncx, const int nco, const int ncw)
{
  parallel_for("Flux", ng, GPU_LAMBDA(const size_t i) {
    dstype param1 = param[0];
    dstype param2 = param[1];
    dstype param3 = param[2];
    dstype param4 = param[3];
    dstype udg1 = udg[0*ng+i];
    dstype udg2 = udg[1*ng+i];
    dstype udg3 = udg[2*ng+i];
    dstype udg4 = udg[3*ng+i];
    dstype udg5 = udg[4*ng+i];
    dstype udg6 = udg[5*ng+i];
    dstype udg7 = udg[6*ng+i];
    dstype udg8 = udg[7*ng+i];
    {
    f[0*ng+i] = param1*udg3+param2*udg1;
    f[1*ng+i] = param1*udg4+param2*udg2;
    f[2*ng+i] = param1*udg5+param3*udg1;
    f[3*ng+i] = param1*udg6+param3*udg2;
    f[4*ng+i] = param1*udg7+param4*udg1;
    f[5*ng+i] = param1*udg8+param4*udg2;
    }
    {
    for (int k = 0; k < 48; ++k) f_udg[k*ng+i] = 0.0;
    f_udg[0*ng+i] = param2;
    f_udg[7*ng+i] = param2;
    f_udg[2*ng+i] = param3;
    f_udg[9*ng+i] = param3;
    f_udg[4*ng+i] = param4;
    f_udg[11*ng+i] = param4;
    f_udg[12*ng+i] = param1;
    f_udg[19*ng+i] = param1;
    f_udg[26*ng+i] = param1;
    f_udg[33*ng+i] = param1;
    f_udg[40*ng+i] = param1;
    f_udg[47*ng+i] = param1;
    }
  });
}
