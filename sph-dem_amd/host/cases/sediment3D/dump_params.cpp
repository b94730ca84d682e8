// dump_params.cpp -- prints the sphdem_params that constants.h builds from
// this case's customConstants.h, as hex bytes (checked against the Python
// binding's parameters of the same case).
#include <cstdio>

#include "customConstants.h"
#include "constants.h"

int main() {
  const sphdem_params P = sphdem_host::paramsFromConstants();
  const unsigned char* b = reinterpret_cast<const unsigned char*>(&P);
  for (size_t i = 0; i < sizeof(P); i++) printf("%02x", b[i]);
  printf("\n");
  return 0;
}
