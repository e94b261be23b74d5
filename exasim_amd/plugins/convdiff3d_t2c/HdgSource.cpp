#include "pdemodel.h"
#include <cmath>

// s = v (the nodal source field, odg)
void HdgSource(dstype* f, dstype* f_udg, dstype* f_wdg, const dstype* xdg, const dstype* udg, const dstype* odg,
               const dstype* wdg, const dstype* uinf, const dstype* param, const dstype time, const int modelnumber,
               const int ng, const int nc, const int ncu, const int nd, const int ncx, const int nco, const int ncw)
{
  parallel_for("Source", ng, GPU_LAMBDA(const size_t i) {
    dstype odg1 = odg[0*ng+i];
    dstype odg2 = odg[1*ng+i];
    {
    f[0*ng+i] = odg1;
    f[1*ng+i] = odg2;
    }
    {
    for (int k = 0; k < 16; ++k) f_udg[k*ng+i] = 0.0;
    }
  });
}
