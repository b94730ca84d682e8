// customOutputBase.h -- the case output plug-in base class
// (customOutputBase.h:10-16 of the reference): run.cpp calls calcOutput at
// each output step with the case's CcustomSim and the file writer.  Like the
// case classes it refers to, it lives in the global namespace.
#pragma once

#include "customConstants.h"
#include "customSim.h"
#include "dataLL.h"
#include "io_data_vtk.h"
#include "vect.h"

class CcustomOutputBase {
 public:
  CcustomOutputBase(CdataLL* _data) : data(_data) {}
  virtual ~CcustomOutputBase() {}
  virtual void calcOutput(int outstep, CcustomSim* custSim, Cio_data_vtk* io) {
    (void)outstep;
    (void)custSim;
    (void)io;
  }

 protected:
  CdataLL* data;
};
