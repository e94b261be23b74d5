#include "pdemodel.h"
#include <cmath>

// f = kappa q; param = (kappa); udg = (u, q_x, q_y)
void HdgFlux(dstype* f, dstype* f_udg, dstype* f_wdg, const dstype* xdg, const dstype* udg, const dstype* odg,
             const dstype* wdg, const dstype* uinf, const dstype* param, const dstype time, const int modelnumber,
             const int ng, const int nc, const int ncu, const int nd, const int ncx, const int nco, const int ncw)
{
  parallel_for("Flux", ng, GPU_LAMBDA(const size_t i) {
    dstype param1 = param[0];
    dstype udg2 = udg[1*ng+i];
    dstype udg3 = udg[2*ng+i];
    {
    f[0*ng+i] = param1*udg2;
    f[1*ng+i] = param1*udg3;
    }
    {
    f_udg[0*ng+i] = 0.0;
    f_udg[1*ng+i] = 0.0;
    f_udg[2*ng+i] = param1;
    f_udg[3*ng+i] = 0.0;
    f_udg[4*ng+i] = 0.0;
    f_udg[5*ng+i] = param1;
    }
  });
}
