// sphIncompress.h -- CsphIncompress(CdataLL*) with start(), middle(), end()
// (sphIncompress.cpp:3-251), each one call into the library; the host
// globals are written before and read back after, so a CcustomSim sees the
// time and custom[] values the reference would show it between the phases.
#pragma once

#include "dataLL.h"

namespace sphdem_host {

class CsphIncompress {
 public:
  explicit CsphIncompress(CdataLL* _data) : data(_data) {}
  void start() { phase(sphdem_step_start); }
  void middle() { phase(sphdem_step_middle); }
  void end() { phase(sphdem_step_end); }

 private:
  void phase(int (*fn)(sphdem_ctx*)) {
    data->pushGlobals();
    data->check(fn(data->context()));
    data->pullGlobals();
  }
  CdataLL* data;
};

}  // namespace sphdem_host
