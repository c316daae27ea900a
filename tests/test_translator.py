"""physisc-hip on the CPU: the committed translations are what the
translator produces, the stencil analysis finds the reference's offsets,
and unsupported constructs are rejected with an error."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHYSISC = os.path.join(ROOT, "physis_amd", "bin", "physisc-hip")
SOURCES = [os.path.join(ROOT, "physis_amd", "programs", f)
           for f in ("diffusion7.c", "himeno.c", "redblack7.c", "udt7.c",
                     "stencil2d.c", "mixed7.c")] + \
          [os.path.join(ROOT, "tests", "programs", "systests.c")]


def translate(src, text=None):
    with tempfile.TemporaryDirectory() as d:
        if text is not None:
            src = os.path.join(d, "prog.c")
            open(src, "w").write(text)
        out = os.path.join(d, "out.hip.cpp")
        p = subprocess.run([PHYSISC, src, "-o", out], capture_output=True,
                           text=True)
        return p.returncode, (open(out).read() if p.returncode == 0 else ""), \
            p.stderr


@pytest.mark.parametrize("src", SOURCES)
def test_committed_translation_is_current(src):
    rc, code, err = translate(src)
    assert rc == 0, err
    committed = open(src[:-2] + ".hip.cpp").read()
    assert code == committed, f"re-run physisc-hip on {src}"


def test_himeno_analysis_finds_19_points():
    _, code, _ = translate(SOURCES[1])
    assert "p0  staged x[-1,1] y[-1,1] z[-1,1] non-periodic" in code
    assert " 18:(0,0,0)" in code
    for g in ("a0", "a1", "a2", "a3", "b0", "b1", "b2", "c0", "c1", "c2",
              "bnd", "wrk1"):
        assert f"//   {g}  direct (centre only)" in code


PROG = """#include "physis/physis.h"
static void k(const int x, const int y, const int z, PSGrid3DFloat a,
              PSGrid3DFloat b, int s) {
  PSGridEmit(b, PSGridGet(a, x, y, %s) + PSGridGetPeriodic(a, x+3, y, z));
}
void run(PSGrid3DFloat a, PSGrid3DFloat b) {
  PSDomain3D d = PSDomain3DNew(0, 8, 0, 8, 0, 8);
  PSStencilRun(PSStencilMap(k, d, a, b, 1), 3);
}
"""


def test_irregular_get_reads_memory():
    rc, code, err = translate(None, PROG % "s")
    assert rc == 0, err
    assert "physis_hip::GetLinear<float>" in code
    assert "x[-3,3]" in code and "periodic" in code


def test_two_dimensional_kernel_runs_in_the_rows_view():
    """A 2-D map over 2-D grids runs on the 2.5D kernel with its grids
    viewed as nx x 1 x ny (y marched as z): offsets (dx, dy) become
    (dx, 0, dy), the stencil sees its 2-D indices and extents, and the
    blocks are one wave.  A 2-D map reading a 1-D grid, and a 1-D map, keep
    the generic kernel."""
    src = """#include "physis/physis.h"
static void k(const int x, const int y, PSGrid2DFloat a, PSGrid2DFloat b) {
  float v = PSGridGet(a, x+1, y) + PSGridGet(a, x-1, y - 1);
  if (y < PSGridDim(a, 1) - 2) v += PSGridGet(a, x, y + 2);
  PSGridEmit(b, v);
}
static void m(const int x, const int y, PSGrid2DFloat a, PSGrid1DFloat c,
              PSGrid2DFloat b) {
  PSGridEmit(b, PSGridGet(a, x, y) * PSGridGet(c, x));
}
static void l(const int x, PSGrid1DFloat c, PSGrid1DFloat e) {
  PSGridEmit(e, PSGridGet(c, x + 1));
}
void run(PSGrid2DFloat a, PSGrid2DFloat b, PSGrid1DFloat c, PSGrid1DFloat e) {
  PSDomain2D d = PSDomain2DNew(1, 7, 0, 8);
  PSDomain1D d1 = PSDomain1DNew(0, 7);
  PSStencilRun(PSStencilMap(k, d, a, b), PSStencilMap(m, d, a, c, b));
  PSStencilRun(PSStencilMap(l, d1, c, e));
}
"""
    rc, code, err = translate(None, src)
    assert rc == 0, err
    assert "Stencil analysis of k (2.5D run kernel over the rows view" in code
    assert "offsets 0:(1,0,0) 1:(-1,0,-1) 2:(0,0,2)" in code
    assert "staged x[-1,1] y[-0,0] z[-1,2] non-periodic" in code
    assert "k(x, z, ctx);" in code
    assert "ctx.a_dim[1] = a.dim[2];" in code  # PSGridDim(a, 1) is ny
    assert "__PSStencilRun_k<VX, 1, -1>), grid, block, 0, stream, " \
        "physis_hip::RowDomain(s.dom), physis_hip::RowView<" \
        "__PSGrid3DFloat_dev>(s.a)" in code
    assert "physis_hip::WindowChunkLimitRows({s.a, s.b}, 3)" in code
    # the y test folds in the interior form: y, the kernel's z, off the
    # first and last rows of a
    assert "z >= 1 && z <= ctx.a_dim[1] - 2" in code
    assert "__PSStencilRunGeneric_m" in code and "__PSStencilRunGeneric_l" in code
    assert "(generic run kernel: a 2-D stencil over grids that are not all " \
        "2-D grids of a primitive type)" in code


def test_three_dimensional_map_reading_lower_dimensional_grids():
    """A 3-D map that reads 1-D and 2-D coefficient grids (the reference's
    test_mixed-dim3 pattern) runs on the 2.5D kernel: the 3-D grid is
    staged, the others are read through bounds-checked gets with their own
    extents, and only 3-D grids decide the vector width and the windows; a
    map emitting a 2-D grid keeps the generic kernel."""
    src = """#include "physis/physis.h"
static void k(const int x, const int y, const int z, PSGrid3DFloat g,
              PSGrid3DFloat g2, PSGrid1DFloat i, PSGrid2DFloat j) {
  PSGridEmit(g2, PSGridGet(g, x - 1, y, z) * PSGridGet(i, x - 1) +
                 PSGridGet(g, x, y, z + 1) * PSGridGet(j, y, z + 1));
}
static void m(const int x, const int y, const int z, PSGrid3DFloat g,
              PSGrid2DFloat j) {
  PSGridEmit(j, PSGridGet(g, x, y, z));
}
void run(PSGrid3DFloat g, PSGrid3DFloat g2, PSGrid1DFloat i, PSGrid2DFloat j) {
  PSDomain3D d = PSDomain3DNew(1, 7, 1, 7, 1, 7);
  PSStencilRun(PSStencilMap(k, d, g, g2, i, j), PSStencilMap(m, d, g, j));
}
"""
    rc, code, err = translate(None, src)
    assert rc == 0, err
    assert "Stencil analysis of k (2.5D run kernel)" in code
    assert "i  read from memory, irregular gets" in code
    assert "physis_hip::GetLinear<float>(__ctx.j_p, __ctx.j_dim, 2, y, z + 1)" in code
    assert "ctx.i_dim[d] = d < 1 ? i.dim[d] : 1;" in code
    assert "ctx.j_dim[d] = d < 2 ? j.dim[d] : 1;" in code
    assert "physis_hip::WindowChunkLimit({s.g, s.g2}, 2)" in code
    assert "(s.g->dim[0] % 4 == 0 && s.g2->dim[0] % 4 == 0) ?" in code
    assert "__builtin_assume((ctx.i_dim[0] & 3) == 0)" not in code
    assert "physis_hip::DevView<__PSGrid1DFloat_dev>(s.i)" in code
    assert ("Stencil analysis of m (generic run kernel: grid j is emitted and "
            "not 3-D)") in code


def test_red_black_maps_visit_one_colour():
    src = """#include "physis/physis.h"
static void k(const int x, const int y, const int z, PSGrid3DFloat g) {
  PSGridEmit(g, PSGridGet(g, x+1, y, z) + PSGridGet(g, x, y, z-1));
}
void run(PSGrid3DFloat g) {
  PSDomain3D d = PSDomain3DNew(1, 7, 1, 7, 1, 7);
  PSStencilRun(PSStencilMapRedBlack(k, d, g), PSStencilMapBlack(k, d, g));
}
"""
    rc, code, err = translate(None, src)
    assert rc == 0, err
    # every get is of the other colour: the 2.5D kernel, the colour in the
    # store mask (the points with x + y + z + rb even)
    assert "(__rb < 0 || ((x + yrow + z + __rb) & 1) == 0)" in code
    assert "__PSStencilRunGeneric_k" not in code
    assert "__PSStencilMapRedBlack_k(d, g)" in code
    assert "__PSStencilMapBlack_k(d, g)" in code
    assert "reads only the other colour" in code and not err
    assert "no (a red-black map not reading its centre)" in code
    # with its centre read, a map of both colours run alone becomes one
    # pass per iteration (red sweep, then black), guarded at run time
    one = src.replace("PSGridGet(g, x+1, y, z) +",
                      "PSGridGet(g, x, y, z) * 0.5f + PSGridGet(g, x+1, y, z) +")
    one = one.replace(", PSStencilMapBlack(k, d, g));", ", 3);")
    rc, code, err = translate(None, one)
    assert rc == 0, err
    assert "red-black: one pass per iteration (red, then black)" in code
    assert "if (iter >= 1 && __PSFusableRB_k(s0)) {" in code
    assert "if (iter % 2) physis_hip::CopyDevice(va.p, vt.p, bytes, stream);" in code
    assert "const __PSDomain gd = physis_hip::WholeDomain(s.g);" in code
    assert "(((xs + i) + yrow + s + 0) & 1) == 0) ? ctx.g_v : a_r" in code
    assert "(((xs + i) + yrow + z + 1) & 1) == 0) ? ctx.g_v : b_r" in code
    # a get of the point's own colour is flagged and keeps the generic
    # kernel, which skips the other colour's points
    rc, code, err = translate(None, src.replace("x+1, y, z", "x+1, y+1, z"))
    assert rc == 0 and "reads its own colour" in err
    assert "if (__rb >= 0 && ((x + y + z + __rb) & 1)) continue;" in code


def test_missing_kernel_is_an_error():
    src = PROG.replace("PSStencilMap(k,", "PSStencilMap(nokernel,") % "z"
    rc, _, err = translate(None, src)
    assert rc != 0 and "nokernel" in err


REF_CASES = "/root/reference/tests/system_tests/test_cases"


@pytest.mark.skipif(not os.path.isdir(REF_CASES),
                    reason="the reference's system tests are not present")
def test_reference_corpus_translates():
    """Every program of the reference's system-test corpus (unchanged) is
    translated, except the features named in the error of the rest."""
    from test_gpu_refcorpus import CORPUS, UNSUPPORTED
    inc = os.path.join(ROOT, "include")
    for name in CORPUS:
        src = os.path.join(REF_CASES, name + ".c")
        if not os.path.exists(src):
            src += "c"  # the C++ test
        with tempfile.TemporaryDirectory() as d:
            p = subprocess.run([PHYSISC, "-I", inc, src, "-o",
                                os.path.join(d, "o.hip.cpp")],
                               capture_output=True, text=True)
        if name in UNSUPPORTED:
            assert p.returncode != 0, name
        else:
            assert p.returncode == 0, (name, p.stderr)


UDT_PROG = """#include "physis/physis.h"
struct Cell {
  float u;
  float m[2][3];
  double w;
};
DeclareGrid3D(Cell, struct Cell);
static void k(const int x, const int y, const int z, PSGrid3DCell g,
              PSGrid3DCell h) {
  int i;
  float s = PSGridGet(g, x + 1, y, z).u + PSGridGet(g, x, y, z).m[1][2];
  for (i = 0; i < 2; ++i) s += PSGridGetPeriodic(g, x, y - 1, z).m[i][0];
  struct Cell c = PSGridGet(g, x, y, z);
  c.u = s;
  PSGridEmit(h, c);
  PSGridEmitUtype(g.w, (double)s);
}
void run(PSGrid3DCell g, PSGrid3DCell h) {
  PSDomain3D d = PSDomain3DNew(1, 7, 1, 7, 1, 7);
  PSStencilRun(PSStencilMap(k, d, g, h));
}
void make(void) {
  PSGrid3DCell g = PSGrid3DCellNew(8, 8, 8);
  PSGridFree(g);
}
"""


def test_user_point_type_lowering():
    rc, code, err = translate(None, UDT_PROG)
    assert rc == 0, err
    # type info as the REF target builds it: scalar size, rank, extents
    assert "{PS_FLOAT, sizeof(float), 2, {2, 3}}" in code
    assert "{PS_DOUBLE, sizeof(double), 0, {0}}" in code
    assert "__PSTypeInfo_Cell = {PS_USER, sizeof(struct Cell), 3" in code
    # one plane per scalar component: member gets index component planes
    assert "physis_hip::UdtLoad<float>(__ctx.g_u.m_m, __ctx.g_u.n, 6, " \
        "((1)) * 3 + (2)" in code
    assert "((i)) * 3 + (0)" in code and "OffsetPeriodic" in code
    # whole-point load / store and a member emit at the current point
    assert "__PSUdtLoad_Cell(__ctx.g_u, physis_hip::OffsetLinear(" in code
    assert "__PSUdtStore_Cell(__ctx.h_u, " in code
    assert "physis_hip::UdtStore<double>(__ctx.g_u.m_w, __ctx.g_u.n, 1, 0, " in code
    # grids of the type are created from the type info
    assert "physis_hip::NewUserGrid(&__PSTypeInfo_Cell, 3, 8, 8, 8)" in code
    assert "generic run kernel: grid g holds user-defined points" in code


SPLIT_PROG = """#include "physis/physis.h"
struct Cplx {
  float r;
  float i;
};
DeclareGrid3D(Cplx, struct Cplx);
struct Arr {
  float p[3][2];
  int tag;
};
DeclareGrid3D(Arr, struct Arr);
static struct Cplx twice(struct Cplx a) {
  struct Cplx b = {2 * a.r, 2 * a.i};
  return b;
}
static void whole(const int x, const int y, const int z, PSGrid3DCplx g1,
                  PSGrid3DCplx g2) {
  struct Cplx a = PSGridGet(g1, x + 1, y, z);
  struct Cplx b = PSGridGet(g1, x, y, z - 1);
  struct Cplx v = {a.r + b.i, a.i - PSGridGet(g1, x, y + 1, z).r};
  PSGridEmit(g2, v);
}
static void parts(const int x, const int y, const int z, PSGrid3DArr g) {
  float s = PSGridGet(g, x - 1, y, z).p[2][1] + PSGridGet(g, x, y, z).p[0][0];
  PSGridEmitUtype(g.p[1][0], s + PSGridGet(g, x, y, z + 1).tag);
}
static void impure(const int x, const int y, const int z, PSGrid3DCplx g1,
                   PSGrid3DCplx g2) {
  PSGridEmit(g2, twice(PSGridGet(g1, x, y, z)));
}
static void looped(const int x, const int y, const int z, PSGrid3DArr g) {
  int i;
  float s = 0;
  for (i = 0; i < 3; ++i) s += PSGridGet(g, x, y, z).p[i][0];
  PSGridEmitUtype(g.p[0][1], s);
}
void run(PSGrid3DCplx g1, PSGrid3DCplx g2, PSGrid3DArr g) {
  PSDomain3D d = PSDomain3DNew(1, 7, 1, 7, 1, 7);
  PSStencilRun(PSStencilMap(whole, d, g1, g2), PSStencilMap(parts, d, g),
               PSStencilMap(impure, d, g1, g2), PSStencilMap(looped, d, g));
}
"""


def test_user_point_grids_split_into_member_grids():
    """A grid of user-defined points touched plane by plane (scalar members,
    literal components of array members, whole points of a struct of
    scalars) runs on the 2.5D kernel over one grid per plane."""
    rc, code, err = translate(None, SPLIT_PROG)
    assert rc == 0, err
    assert "Stencil analysis of whole (2.5D run kernel)" in code
    # a local point initialised by a whole read reads only the members it
    # uses: b uses b.i alone, so g1__r is not read at (0, 0, -1)
    assert "g1__r  staged x[-1,1] y[-1,1] z[0,0] non-periodic" in code
    assert "g1__i  staged x[-1,1] y[-0,0] z[-1,0] non-periodic" in code
    assert "struct Cplx b = __PSUdtMake_Cplx(0, __ctx.g1__i[1]);" in code
    # a whole point read member by member, a whole point emitted likewise
    assert "__PSUdtMake_Cplx(__ctx.g1__r[" in code
    assert "(__ctx.g2__r_v = ((v).r), __ctx.g2__r_e = true)" in code
    assert "(__ctx.g2__i_v = ((v).i), __ctx.g2__i_e = true)" in code
    # component planes: p[2][1] is component 5 of member 0, p[1][0] is 2
    assert "Stencil analysis of parts (2.5D run kernel)" in code
    assert "physis_hip::MemberView<__PSGrid3DFloat_dev>(s.g, 0, 5)" in code
    assert "physis_hip::MemberView<__PSGrid3DFloat_dev>(s.g, 0, 2)" in code
    assert "physis_hip::MemberView<__PSGrid3DInt_dev>(s.g, 1, 0)" in code
    assert "physis_hip::CheckUserGrid3D(s.g);" in code
    # a point passed to a function, a component under a loop index: whole
    # points through the generic kernel (g1 of impure is still read member
    # by member; the emitted value, a call, is not repeated per member)
    assert ("Stencil analysis of impure (generic run kernel: grid g2 holds "
            "user-defined points)") in code
    assert "__PSStencilRunGeneric_impure<-1>), grid, dim3(256), 0, stream, " \
        "s.dom, __PSUdtViewOf_Cplx(s.g2), physis_hip::MemberView<" \
        "__PSGrid3DFloat_dev>(s.g1, 0, 0)" in code
    assert ("Stencil analysis of looped (generic run kernel: grid g holds "
            "user-defined points)") in code


def test_user_point_scalar_member_indexed_is_not_split():
    """An index applied to a scalar member of a get is not a plane: the
    grid is not split, and the generic lowering reports the error."""
    bad = SPLIT_PROG.replace("PSGridGet(g, x, y, z + 1).tag",
                             "PSGridGet(g, x, y, z + 1).tag[0]")
    rc, _, err = translate(None, bad)
    assert rc != 0 and "tag" in err


def test_user_point_member_must_be_indexed_to_a_scalar():
    bad = UDT_PROG.replace("PSGridEmitUtype(g.w, (double)s);",
                           "PSGridEmitUtype(g.m[0], s);")
    rc, _, err = translate(None, bad)
    assert rc != 0 and "indexed down to a scalar" in err


def test_interior_split_only_for_kernels_comparing_y_or_z():
    d7 = open(SOURCES[0][:-2] + ".hip.cpp").read()
    him = open(SOURCES[1][:-2] + ".hip.cpp").read()
    # the clamped diffusion compares y and z with the grid's first and last
    # planes: its wave steps have an interior form
    assert "__builtin_assume(yrow != ctx.g1_dim[1] - 1)" in d7
    assert "__pts(physis_hip::Flag<true>{})" in d7
    # Himeno never compares an index
    assert "Flag<true>" not in him
    # x-alignment facts of the VX = 4 form
    assert "__builtin_assume((xs & (VX - 1)) == 0);" in d7


def _translate_with_config(conf_text, src):
    with tempfile.TemporaryDirectory() as d:
        conf = os.path.join(d, "opt.conf")
        open(conf, "w").write(conf_text)
        out = os.path.join(d, "out.hip.cpp")
        p = subprocess.run([PHYSISC, "--config", conf, src, "-o", out],
                           capture_output=True, text=True)
        return p.returncode, (open(out).read() if p.returncode == 0 else ""), \
            p.stderr


def test_reference_configuration_files_are_accepted():
    # the reference's own configuration files (examples/diffusion-benchmark/
    # opt.conf, examples/himeno/autotune.conf): optimizer flags, Lua tables
    conf = """-- optimizer flags of the reference
OPT_OFFSET_COMP = true
OPT_LOOP_PEELING = {false, true}
OPT_REGISTER_BLOCKING = true
CUDA_BLOCK_SIZE = {{32, 4, 1}, {64, 4, 1}}
CUDA_KERNEL_ERROR_CHECK = true
HIP_TWO_SWEEP = false
"""
    rc, code, err = _translate_with_config(conf, SOURCES[0])
    assert rc == 0 and "unknown configuration key" not in err, err
    assert "Flag<true>" not in code             # OPT_LOOP_PEELING = false
    assert "__PSStencilRun2_" not in code       # HIP_TWO_SWEEP = false
    assert "__PSCheckHipError" in code          # CUDA_KERNEL_ERROR_CHECK
    rc, _, err = _translate_with_config("NOT_A_KEY = 1\n", SOURCES[0])
    assert rc == 0 and "unknown configuration key NOT_A_KEY" in err


REF_EXAMPLES = "/root/reference/examples"


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES),
                    reason="the reference's examples are not present")
@pytest.mark.parametrize("rel", [
    "test_double_buffering.c", "test_global_variable.c", "test_set.c",
    "test_staggered_grid.c", "diffusion-benchmark/diffusion3d_physis.c",
    "himeno/himenobmtxpa_physis.c", "copy/copy3d.c"])
def test_reference_examples_translate(rel):
    """The reference's example and benchmark programs, unchanged."""
    src = os.path.join(REF_EXAMPLES, rel)
    with tempfile.TemporaryDirectory() as d:
        p = subprocess.run([PHYSISC, "-I", os.path.join(ROOT, "include"),
                            "-I", os.path.dirname(src), src, "-o",
                            os.path.join(d, "o.hip.cpp")],
                           capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_host_get_reads_one_element():
    src = """#include "physis/physis.h"
float first(PSGrid3DFloat g) { return PSGridGet(g, 0, 1, 2); }
"""
    rc, code, err = translate(None, src)
    assert rc == 0, err
    assert "physis_hip::GetElem<float>(g, 0, 1, 2)" in code
    rc, _, err = translate(None, src.replace("PSGridGet(", "PSGridGetPeriodic("))
    assert rc != 0 and "not allowed in host code" in err


CXX_PROG = """#include <iostream>
#include "physis/physis.h"
using namespace std;
template <typename T> T twice(T v) { return v + v; }
namespace util { inline int one() { return 1; } }
class Scale {
 public:
  Scale(int s) { s_ = s; }
  int operator()() { return s_; }
 private:
  int s_;
};
void copy(const int x, const int y, const int z, PSGrid3DFloat a,
          PSGrid3DFloat b) {
  PSGridEmit(b, PSGridGet(a, x, y, z));
}
void run(Scale &s, PSGrid3DFloat a, PSGrid3DFloat b) {
  float k = (float)s();
  PSDomain3D d = PSDomain3DNew(0, 4, 0, 4, 0, 4);
  PSStencilRun(PSStencilMap(copy, d, a, b));
}
int main(int argc, char *argv[]) {
  PSInit(&argc, &argv, 3, 4, 4, 4);
  Scale s(2);
  return s() - 2;
}
"""


def test_cplusplus_declarations_are_kept():
    """C++ declarations (a class, a using-directive) pass through unchanged
    and outside the C-linkage block; reference parameters and direct
    initialisation "T v(args)" are parsed (the reference's test_cplusplus.cc
    uses all of them)."""
    rc, code, err = translate(None, CXX_PROG)
    assert rc == 0, err
    cls = code.index("class Scale")
    assert code.index("using namespace std;") < cls
    assert 'extern "C" {' not in code[:cls]
    assert "void run(Scale & s, PSGrid3DFloat a, PSGrid3DFloat b)" in code
    assert "Scale s(2);" in code and "(float)s()" in code
    # a function template and a namespace end with their closing brace
    assert "template <typename T> T twice(T v) { return v + v; }" in code
    assert code.index("namespace util") < code.index("class Scale")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "cxx.hip.cpp")
        open(src, "w").write(code)
        p = subprocess.run(["/opt/rocm/bin/hipcc", "-x", "hip",
                            "--offload-arch=gfx950", "-std=c++17",
                            "-fsyntax-only", "-w",
                            "-I" + os.path.join(ROOT, "include"), src],
                           capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-3000:]


@pytest.mark.parametrize("body", [
    # shadowed in a block
    "int xx = x; { int x = xx + 2; v = PSGridGet(a, x - 1, y, z); }",
    # assigned
    "int xx = x; xx = xx; v = PSGridGet(a, x + 1, y, z); x = x + 0;",
    # incremented
    "z++; v = PSGridGet(a, x, y, z - 1);",
    # a loop variable with an index name
    "v = 0; for (int y = 0; y < 2; ++y) v += PSGridGet(a, x, y, z);",
])
def test_rebound_index_makes_gets_irregular(body):
    src = """#include "physis/physis.h"
static void k(int x, int y, int z, PSGrid3DFloat a, PSGrid3DFloat b) {
  float v;
  %s
  PSGridEmit(b, v);
}
void run(PSGrid3DFloat a, PSGrid3DFloat b) {
  PSDomain3D d = PSDomain3DNew(0, 8, 0, 8, 0, 8);
  PSStencilRun(PSStencilMap(k, d, a, b), PSStencilMap(k, d, b, a), 2);
}
""" % body
    rc, code, err = translate(None, src)
    assert rc == 0, err
    assert "a  read from memory, irregular gets" in code
    assert "ReadRow" not in code
    assert "two-sweep fusion of map pairs: no (an index parameter is " \
           "rebound)" in code


def test_unrebound_index_stays_regular():
    src = """#include "physis/physis.h"
static void k(int x, int y, int z, PSGrid3DFloat a, PSGrid3DFloat b) {
  int xx = x + 1;
  PSGridEmit(b, PSGridGet(a, x + 1, y, z) + xx);
}
void run(PSGrid3DFloat a, PSGrid3DFloat b) {
  PSDomain3D d = PSDomain3DNew(0, 8, 0, 8, 0, 8);
  PSStencilRun(PSStencilMap(k, d, a, b));
}
"""
    rc, code, err = translate(None, src)
    assert rc == 0, err
    assert "a  staged x[-1,1]" in code
