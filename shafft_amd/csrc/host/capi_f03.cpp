// Fortran-facing companions of the configuration objects: heap allocation
// of the C structs, wrappers that take the Fortran communicator handle, and
// the per-field accessors the Fortran module's derived types sync through
// (src/f03/shafftf03.h:46-146).  The handle must hold 0, the one-rank
// communicator; accessors on a null config return 0 / do nothing.
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "../../../include/shafft_amd.h"

namespace {

void copy_name(const char* src, size_t len, char* out, size_t maxLen) {
  if (!out || maxLen == 0) return;
  const size_t n = std::min(len, maxLen);
  if (src && n) std::memcpy(out, src, n);
  if (n < maxLen) out[n] = '\0';
}

template <typename T>
void copy_array(const T* src, int n, T* out) {
  if (src && out && n > 0) std::memcpy(out, src, (size_t)n * sizeof(T));
}

// Fortran handle -> the C one-rank handle; anything but 0 is handed on as is,
// so the C entry point reports ERR_INVALID_COMM.
shafft_comm_t from_fortran(const shafft_fint_t* h) { return h ? (shafft_comm_t)*h : (shafft_comm_t)-1; }

}  // namespace

extern "C" {

shafft_nd_config_t* shafftConfigNDAllocf03(void) {
  return static_cast<shafft_nd_config_t*>(std::calloc(1, sizeof(shafft_nd_config_t)));
}
void shafftConfigNDFreef03(shafft_nd_config_t* cfg) { std::free(cfg); }
shafft_1d_config_t* shafftConfig1DAllocf03(void) {
  return static_cast<shafft_1d_config_t*>(std::calloc(1, sizeof(shafft_1d_config_t)));
}
void shafftConfig1DFreef03(shafft_1d_config_t* cfg) { std::free(cfg); }

int shafftConfigNDInitf03(shafft_nd_config_t* cfg, int ndim, const size_t* globalShape,
                          shafft_t precision, const int* commDims, int hintNda,
                          shafft_decomposition_strategy_t strategy,
                          shafft_transform_layout_t outputPolicy, size_t memLimit,
                          shafft_fint_t* fComm) {
  return shafftConfigNDInit(cfg, ndim, globalShape, precision, commDims, hintNda, strategy,
                            outputPolicy, memLimit, from_fortran(fComm));
}
int shafftConfigNDResolvef03(shafft_nd_config_t* cfg) { return shafftConfigNDResolve(cfg); }
int shafftNDInitFromConfigf03(void* planPtr, shafft_nd_config_t* cfg) {
  return shafftNDInitFromConfig(planPtr, cfg);
}
int shafftConfig1DInitf03(shafft_1d_config_t* cfg, size_t globalSize, shafft_t precision,
                          shafft_fint_t* fComm) {
  return shafftConfig1DInit(cfg, globalSize, precision, from_fortran(fComm));
}
int shafftConfig1DResolvef03(shafft_1d_config_t* cfg) { return shafftConfig1DResolve(cfg); }
int shafft1DInitFromConfigf03(void* planPtr, shafft_1d_config_t* cfg) {
  return shafft1DInitFromConfig(planPtr, cfg);
}
int shafftGetCommunicatorf03(void* planPtr, shafft_fint_t* outFint) {
  shafft_comm_t c = 0;
  const int rc = shafftGetCommunicator(planPtr, outFint ? &c : nullptr);
  if (rc == 0) *outFint = (shafft_fint_t)c;
  return rc;
}

// Each accessor returns its field, or the reference's value for a null
// config (src/f03/shafftf03.cpp: 0, and -1 for activeRank).
#define SA_GET_OR(T, CFG, NAME, EXPR, NONE) \
  T shafftConfig##NAME(const CFG* cfg) { return cfg ? (T)(EXPR) : (T)(NONE); }
#define SA_GET(T, CFG, NAME, EXPR) SA_GET_OR(T, CFG, NAME, EXPR, 0)

SA_GET(int, shafft_nd_config_t, NDGetNdim, cfg->ndim)
SA_GET(int, shafft_nd_config_t, NDGetPrecision, cfg->precision)
SA_GET(size_t, shafft_nd_config_t, NDGetAllocElements, cfg->allocElements)
SA_GET(int, shafft_nd_config_t, NDGetIsActive, cfg->isActive)
SA_GET_OR(int, shafft_nd_config_t, NDGetActiveRank, cfg->activeRank, -1)
SA_GET(int, shafft_nd_config_t, NDGetActiveSize, cfg->activeSize)
SA_GET(int, shafft_nd_config_t, NDGetNda, cfg->nda)
SA_GET(int, shafft_nd_config_t, NDGetCommSize, cfg->commSize)
SA_GET(int, shafft_nd_config_t, NDGetStatus, cfg->status)
SA_GET(int, shafft_nd_config_t, NDGetOutputPolicy, cfg->outputPolicy)
SA_GET(int, shafft_nd_config_t, NDGetStrategy, cfg->strategy)
SA_GET(size_t, shafft_nd_config_t, NDGetMemLimit, cfg->memLimit)
SA_GET(int, shafft_nd_config_t, NDGetHintNda, cfg->hintNda)
SA_GET(int, shafft_nd_config_t, NDGetFlags, cfg->flags)
SA_GET(shafft_fint_t, shafft_nd_config_t, NDGetWorldComm, cfg->worldComm)
SA_GET(shafft_fint_t, shafft_nd_config_t, NDGetActiveComm, cfg->activeComm)
SA_GET(int, shafft_nd_config_t, NDGetNodeId, cfg->nodeId)
SA_GET(int, shafft_nd_config_t, NDGetNodeCount, cfg->nodeCount)
SA_GET(size_t, shafft_nd_config_t, NDGetHostnameLen, cfg->hostnameLen)
SA_GET(size_t, shafft_nd_config_t, NDGetDeviceNameLen, cfg->deviceNameLen)
SA_GET(size_t, shafft_1d_config_t, 1DGetGlobalSize, cfg->globalSize)
SA_GET(int, shafft_1d_config_t, 1DGetPrecision, cfg->precision)
SA_GET(size_t, shafft_1d_config_t, 1DGetAllocElements, cfg->allocElements)
SA_GET(int, shafft_1d_config_t, 1DGetIsActive, cfg->isActive)
SA_GET_OR(int, shafft_1d_config_t, 1DGetActiveRank, cfg->activeRank, -1)
SA_GET(int, shafft_1d_config_t, 1DGetActiveSize, cfg->activeSize)
SA_GET(int, shafft_1d_config_t, 1DGetStatus, cfg->status)
SA_GET(int, shafft_1d_config_t, 1DGetFlags, cfg->flags)
SA_GET(shafft_fint_t, shafft_1d_config_t, 1DGetWorldComm, cfg->worldComm)
SA_GET(shafft_fint_t, shafft_1d_config_t, 1DGetActiveComm, cfg->activeComm)
SA_GET(int, shafft_1d_config_t, 1DGetNodeId, cfg->nodeId)
SA_GET(int, shafft_1d_config_t, 1DGetNodeCount, cfg->nodeCount)
SA_GET(size_t, shafft_1d_config_t, 1DGetHostnameLen, cfg->hostnameLen)
SA_GET(size_t, shafft_1d_config_t, 1DGetDeviceNameLen, cfg->deviceNameLen)
SA_GET(size_t, shafft_1d_config_t, 1DGetInitialLocalSize, cfg->initial.localSize)
SA_GET(size_t, shafft_1d_config_t, 1DGetInitialLocalStart, cfg->initial.localStart)
SA_GET(size_t, shafft_1d_config_t, 1DGetOutputLocalSize, cfg->output.localSize)
SA_GET(size_t, shafft_1d_config_t, 1DGetOutputLocalStart, cfg->output.localStart)
#undef SA_GET

void shafftConfigNDGetHostname(const shafft_nd_config_t* cfg, char* out, size_t maxLen) {
  if (cfg) copy_name(cfg->hostname, cfg->hostnameLen, out, maxLen);
}
void shafftConfigNDGetDeviceName(const shafft_nd_config_t* cfg, char* out, size_t maxLen) {
  if (cfg) copy_name(cfg->deviceName, cfg->deviceNameLen, out, maxLen);
}
void shafftConfig1DGetHostname(const shafft_1d_config_t* cfg, char* out, size_t maxLen) {
  if (cfg) copy_name(cfg->hostname, cfg->hostnameLen, out, maxLen);
}
void shafftConfig1DGetDeviceName(const shafft_1d_config_t* cfg, char* out, size_t maxLen) {
  if (cfg) copy_name(cfg->deviceName, cfg->deviceNameLen, out, maxLen);
}

void shafftConfigNDGetGlobalShape(const shafft_nd_config_t* cfg, size_t* out) {
  if (cfg) copy_array(cfg->globalShape, cfg->ndim, out);
}
void shafftConfigNDGetCommDims(const shafft_nd_config_t* cfg, int* out) {
  if (cfg) copy_array(cfg->commDims, cfg->ndim, out);
}
void shafftConfigNDGetHintCommDims(const shafft_nd_config_t* cfg, int* out) {
  if (cfg) copy_array(cfg->hintCommDims, cfg->ndim, out);
}
void shafftConfigNDGetInitialSubsize(const shafft_nd_config_t* cfg, size_t* out) {
  if (cfg) copy_array(cfg->initial.subsize, cfg->ndim, out);
}
void shafftConfigNDGetInitialOffset(const shafft_nd_config_t* cfg, size_t* out) {
  if (cfg) copy_array(cfg->initial.offset, cfg->ndim, out);
}
void shafftConfigNDGetOutputSubsize(const shafft_nd_config_t* cfg, size_t* out) {
  if (cfg) copy_array(cfg->output.subsize, cfg->ndim, out);
}
void shafftConfigNDGetOutputOffset(const shafft_nd_config_t* cfg, size_t* out) {
  if (cfg) copy_array(cfg->output.offset, cfg->ndim, out);
}

void shafftConfigNDSetOutputPolicy(shafft_nd_config_t* cfg, int pol) {
  if (cfg) cfg->outputPolicy = (shafft_transform_layout_t)pol;
}
void shafftConfigNDSetStrategy(shafft_nd_config_t* cfg, int s) {
  if (cfg) cfg->strategy = (shafft_decomposition_strategy_t)s;
}
void shafftConfigNDSetMemLimit(shafft_nd_config_t* cfg, size_t limit) {
  if (cfg) cfg->memLimit = limit;
}
void shafftConfigNDSetHintNda(shafft_nd_config_t* cfg, int nda) {
  if (cfg) cfg->hintNda = nda;
}
void shafftConfigNDSetHintCommDims(shafft_nd_config_t* cfg, const int* dims) {
  if (cfg && dims && cfg->hintCommDims)
    std::memcpy(cfg->hintCommDims, dims, (size_t)cfg->ndim * sizeof(int));
}

}  // extern "C"
