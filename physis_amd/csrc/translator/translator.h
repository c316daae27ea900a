// physisc-hip: translation of Physis programs for the HIP target.
//
// Responsibilities mirror the reference's target plug-in:
//   * stencil analysis (translator/stencil_analysis.cc:64-143,
//     stencil_range.h:136-176): per grid parameter of a stencil function,
//     the constant offsets of its gets, halo widths and periodicity;
//   * HIPTranslator (cf. CUDATranslator, cuda_translator.h:12-57): rewrites
//     PS* intrinsics in host code and emits the program's functions;
//   * HIPRuntimeBuilder (cf. CUDARuntimeBuilder, cuda_runtime_builder.cc:
//     1099-1641): emits stencil structs, map functions, the device form of
//     each stencil function, its run kernel and launcher, and the host run
//     function of each PSStencilRun call site.
#ifndef PHYSIS_AMD_TRANSLATOR_TRANSLATOR_H_
#define PHYSIS_AMD_TRANSLATOR_TRANSLATOR_H_

#include <array>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "ast.h"
#include "parser.h"

namespace physisc {

struct Config {
  // Launch shapes of the 2.5D run kernel (waves per block, prefetch depth).
  int block_waves = 0;  // 0: chosen from the LDS budget
  int prefetch = 0;     // 0: chosen from the register estimate (2 or 4)
  bool kernel_error_check = false;  // HIP_KERNEL_ERROR_CHECK
  bool trace = false;               // TRACE_KERNEL
  bool two_sweep = true;            // HIP_TWO_SWEEP: fuse map pairs
  bool interior_split = true;       // OPT_LOOP_PEELING: interior wave steps
};

// One grid parameter of a stencil function.
struct GridParam {
  std::string name;
  int dims = 3;
  std::string elem;    // C element type: float double int long
  std::string ps_type; // PS_FLOAT ...
  std::string dev_view;  // __PSGrid3DFloat_dev ...
  // analysis
  bool read = false, emitted = false;
  std::vector<std::array<int, 3>> offsets;  // regular gets, first-seen order
  std::vector<bool> offset_periodic;
  bool irregular = false;   // some get is not index+constant
  bool periodic = false;    // staged tile wraps
  int hx = 0, hy = 0, zl = 0, zr = 0;
  enum Mode { kNone, kStaged, kDirect } mode = kNone;
  // a grid of user-defined points (DeclareGrid<n>D(uname, elem)): stored as
  // one plane per scalar member component, accessed through the view
  // __PSUdtView_<uname>
  bool user = false;
  std::string uname;
  const std::vector<StructMember> *members = nullptr;
  // A user grid whose stencil reads and emits it only plane by plane (a
  // scalar member, a literal component of an array member, or whole points
  // of a struct of scalars member by member) is split: split = true on it,
  // and each plane it uses becomes a grid of its own (member_of = the user
  // grid's index, member_index = the member's, member_comp = the component
  // of an array member, row-major).
  bool split = false;
  int member_of = -1, member_index = -1, member_comp = 0;
};

struct ScalarParam {
  std::string type, name;
};

struct KernelInfo {
  const Function *fn = nullptr;
  std::string name;
  int ndims = 3;
  std::vector<std::string> index;  // x y z parameter names
  std::vector<GridParam> grids;
  std::vector<ScalarParam> scalars;
  // parameter order: 'g' grid i / 's' scalar i
  std::vector<std::pair<char, int>> order;
  std::string ret_type = "void";
  bool fast = false;          // 2.5D template path
  // A 2-D kernel on the 2.5D path: its grids are viewed as nx x 1 x ny 3-D
  // grids (the same linear offsets), so the kernel marches y as z with
  // one-wave blocks; offsets are stored as (dx, 0, dy).
  bool rows2d = false;
  // an index parameter is redeclared, assigned, incremented or has its
  // address taken in the body: every get is treated as irregular
  bool index_rebound = false;
  std::string slow_reason;    // why not
  int hy = 0;                 // common y halo of staged grids
  bool used_plain = false, used_reduce = false;
  // used in a red-black map (PSStencilMapRed/Black/RedBlack): the run
  // kernel visits one colour, x + y + z + rb even (rb 0 red, 1 black)
  bool used_rb = false;
  std::string rb_note;  // whether the in-place colour sweep matches REF order
  // two-sweep fusion (temporal blocking) of map pairs k(a->b), k(b->a)
  bool fusable = false;
  // the pair fuses over a domain strictly inside the grid (points outside
  // it keep their values), rather than over the whole grid
  bool fuse_domain = false;
  // a red-black map run as one pass per iteration (red sweep, then black)
  bool fuse_rb = false;
  int fuse_in = -1, fuse_out = -1;  // grid indices of the read and emitted grids
  std::string fuse_why;             // why not, for the analysis comment
  // the body compares index parameter y / z (axis 1 / 2) with something:
  // run kernels then split each wave step into an interior form (y, z
  // known to lie off the grid's first and last planes) and a general one
  bool compares_axis[3] = {false, false, false};
  int GridIndex(const std::string &n) const {
    for (size_t i = 0; i < grids.size(); ++i)
      if (grids[i].name == n) return (int)i;
    return -1;
  }
  const GridParam *Grid(const std::string &n) const {
    int i = GridIndex(n);
    return i < 0 ? nullptr : &grids[i];
  }
};

// A PSStencilRun call site: the kernels of its maps, in order.
struct RunSite {
  int id;
  std::vector<std::string> kernels;
};

class Translator {
 public:
  Translator(Program &prog, const Config &cfg,
             const std::vector<std::pair<std::string, std::string>> &globals);
  std::string Translate();

 private:
  // analysis
  void FindKernels();
  void FindKernelsIn(const Stmt &s);
  void FindKernelsIn(const Expr &e);
  void AnalyzeKernel(KernelInfo &k);
  void AnalyzeGets(KernelInfo &k, const Expr &e);
  void AnalyzeGets(KernelInfo &k, const Stmt &s);
  bool RegularIndex(const KernelInfo &k, const Expr &e, int axis, int *off);
  void Decide(KernelInfo &k);
  bool IsUnusedStencil(const Function &f) const;
  void AnalyzeFusion(KernelInfo &k);

  // device code
  void RewriteKernelExpr(const KernelInfo &k, ExprP &e);
  void RewriteKernelStmt(const KernelInfo &k, Stmt &s);
  std::string EmitKernel(KernelInfo &k);
  std::string EmitStencilStruct(const KernelInfo &k);
  std::string EmitContext(const KernelInfo &k);
  std::string EmitDeviceFunction(KernelInfo &k);
  std::string EmitFastRunKernel(const KernelInfo &k);
  std::string EmitGenericRunKernel(const KernelInfo &k);
  std::string EmitLauncher(const KernelInfo &k);
  std::string EmitFusedRunKernel(const KernelInfo &k);
  std::string EmitFusedDomainKernel(const KernelInfo &k);
  // the run-function branch that runs a fusable pair as two-sweep passes
  std::string EmitFusedRun(const KernelInfo &k);
  // Interior split of a wave step at rows y, plane z: the wave-uniform
  // condition and the facts assumed under it ("" when the body compares
  // neither y nor z).
  // only >= 0: the extents of grid k.grids[only] stand for every grid's
  // (two-sweep passes, whose grids have equal extents).
  std::string InteriorCond(const KernelInfo &k, const std::string &y,
                           const std::string &z, int only = -1) const;
  std::string InteriorAssume(const KernelInfo &k, const std::string &y,
                             const std::string &z, int only = -1) const;
  // the lambda-wrapped element loop `body` called in its two forms
  std::string SplitInterior(const KernelInfo &k, const std::string &y,
                            const std::string &z, const std::string &body,
                            const std::string &ind, int only = -1) const;
  // VGPRs per lane of the run kernel at VX = 4 with pd staged planes in
  // flight (a rough estimate), and the prefetch depth used for k.
  int RegEstimate(const KernelInfo &k, int pd) const;
  int PrefetchFor(const KernelInfo &k) const;
  std::string EmitRunFunction(const RunSite &r);
  // type info, device view and load/store helpers of a user point type
  std::string EmitUserType(const std::string &uname);
  // PSGridGet(g, ...).member[i]... on a grid of user-defined points
  bool RewriteUserMemberGet(const KernelInfo &k, ExprP &e);
  std::string UserOffset(const KernelInfo &k, const GridParam &g,
                         const Expr &call);
  std::string UserPointOffset(const KernelInfo &k, const GridParam &g) const;
  std::string UserComponent(const KernelInfo &k, const GridParam &g,
                            const StructMember &m,
                            std::vector<ExprP *> &idx);

  // host code
  void RewriteHostExpr(ExprP &e, const Function &f);
  void RewriteHostStmt(Stmt &s, const Function &f);
  std::string GridElemOf(const Expr &g, const Function &f) const;
  ExprP MapCall(Expr &call, const Function &f);

  Program &prog_;
  Config cfg_;
  std::map<std::string, std::string> grid_vars_;  // global name -> type
  std::map<std::string, std::map<std::string, std::string>> local_grids_;
  std::map<std::string, KernelInfo> kernels_;
  std::vector<std::string> kernel_order_;
  std::set<std::string> device_helpers_;
  std::vector<RunSite> runs_;
  std::map<const Function *, std::vector<int>> runs_in_;
};

}  // namespace physisc

#endif  // PHYSIS_AMD_TRANSLATOR_TRANSLATOR_H_
