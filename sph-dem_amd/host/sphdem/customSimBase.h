// customSimBase.h -- the case plug-in base class (customSimBase.h:8-22 of the
// reference): a CcustomSim overrides the four hooks run.cpp calls around the
// half-steps.
#pragma once

#include "dataLL.h"

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

 protected:
  double time;
  CdataLL* data;
};

}  // namespace sphdem_host
