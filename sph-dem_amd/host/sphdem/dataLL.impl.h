// dataLL.impl.h -- the reference keeps CdataLL's template bodies here
// (dataLL.impl.h:16-477); in the drop-in they are in dataLL.h, which this
// includes so that case code including either compiles.
#pragma once

#include "dataLL.h"
