// misc.h -- the conversions between the artificial-viscosity coefficient and
// a kinematic viscosity of the Wendland kernel (Nmisc, misc.h:38-46 of the
// reference).
#pragma once

#include "vect.h"

namespace Nmisc {
inline double alphaToVisc(const double alpha, const double h, const double spsound) {
  return (15.0 / 112.0) * alpha * h * spsound;
}
inline double viscToAlpha(const double visc, const double h, const double spsound) {
  return visc * (112.0 / 15.0) / (spsound * h);
}
inline double viscToH(const double visc, const double alpha, const double spsound) {
  return visc * (112.0 / 15.0) / (spsound * alpha);
}
}  // namespace Nmisc
