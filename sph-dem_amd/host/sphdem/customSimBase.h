// customSimBase.h -- the case plug-in base class (customSimBase.h:8-22 of the
// reference): a CcustomSim overrides the four hooks run.cpp calls around the
// half-steps, and may filter velocities and positions (vFilter, rFilter).
#pragma once

#include "customConstants.h"
#include "dataLL.h"
#include "vect.h"

namespace sphdem_host {

class CcustomSimBase {
 public:
  CcustomSimBase(CdataLL* _data, double _time) : time(_time), data(_data) {}
  virtual ~CcustomSimBase() {}
  double getTime() { return time; }
  virtual void beforeStart(double newTime) { time = newTime; }
  virtual void beforeMiddle(double newTime) { time = newTime; }
  virtual void beforeEnd(double newTime) { time = newTime; }
  virtual void afterEnd(double newTime) { time = newTime; }
  virtual vect vFilter(const vect vin, const vect rin) {
    (void)rin;
    return vin;
  }
  virtual vect rFilter(const vect rin) { return rin; }

 protected:
  double time;
  CdataLL* data;
};

}  // namespace sphdem_host

using sphdem_host::CcustomSimBase;
