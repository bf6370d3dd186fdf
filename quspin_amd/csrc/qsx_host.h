// Host-side objects shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "qsx_internal.h"

struct qsx_core {
  CoreDev h;                 // host copy of the device descriptor
  CoreDev* d = nullptr;      // device copy
  double2* phase = nullptr;  // character chi(code), G entries
  std::vector<double2> phase_h;
  // particle sector recorded by make_basis (for LOOKUP_RANK)
  RankDev rank_h;
  RankDev* rank_d = nullptr;
  int rank_ok = 0;           // 1 when the basis is a full single sector
  int sector_ok = 0;         // 1 when make_basis enumerated one sector (rank_h describes it)
  int64_t sector_total = 0;  // states in that sector (before symmetry reduction)
};

struct qsx_oplist {
  int n_off = 0, n_diag = 0;
  TermDev* off = nullptr;    // device, application-order ops
  TermDev* diag = nullptr;
  std::vector<TermDev> off_h, diag_h;
  int complex_me = 0;        // any complex coupling or 'y'
};

void qsx_set_error(const std::string& msg);

// Captured launch sequences (hipGraphExec) hold kernel arguments by value.
// An object whose fields enter those arguments (an H.v context, a factorised
// operator) carries an owner id from qsx_owner_new(); captures are keyed on
// the id, and qsx_graph_drop_owner(id) destroys them when the object changes
// or is destroyed, so a later object at a recycled address is captured anew.
uint64_t qsx_owner_new();
void qsx_graph_drop_owner(uint64_t owner);
// One operator string compiled for the kernels (hv.hip): characters applied
// right to left, hc = 1 for the Hermitian conjugate; n_y counts 'y'.
int qsx_compile_term(const qsx_core* c, const char* opstr, const int32_t* indx, int n_op, double jr, double ji, int hc,
                     TermDev& T, int& n_y);
// owner id of an H.v context (hv.hip)
struct qsx_hv_ctx;
uint64_t qsx_hv_ctx_owner(const qsx_hv_ctx* X);

#define QSX_CHECK_HIP(expr)                                               \
  do {                                                                    \
    hipError_t _e = (expr);                                               \
    if (_e != hipSuccess) {                                               \
      qsx_set_error(std::string("HIP error: ") + hipGetErrorString(_e) + \
                    " at " #expr);                                        \
      return QSX_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

#define QSX_LAUNCH_CHECK()                                                \
  do {                                                                    \
    hipError_t _e = hipGetLastError();                                    \
    if (_e != hipSuccess) {                                               \
      qsx_set_error(std::string("kernel launch failed: ") +              \
                    hipGetErrorString(_e));                               \
      return QSX_ERR_HIP;                                                 \
    }                                                                     \
  } while (0)

enum {
  QSX_OK = 0,
  QSX_ERR_ARG = -1,
  QSX_ERR_HIP = -2,
  QSX_ERR_OVERFLOW = -3,
  QSX_ERR_COMPLEX = -4,
  QSX_ERR_UNSUPPORTED = -5,
};

enum { QSX_F32 = 0, QSX_F64 = 1, QSX_C64 = 2, QSX_C128 = 3 };

static inline int64_t qsx_grid(int64_t n, int block) { return (n + block - 1) / block; }
