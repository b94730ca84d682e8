// sphConstants.h -- the kernel support radius in units of h (sphConstants.h:4-8
// of the reference; the Wendland kernel of the library has support 2h).
#pragma once

#ifndef KERNAL_RADIUS
#define KERNAL_RADIUS 2.0
#endif
