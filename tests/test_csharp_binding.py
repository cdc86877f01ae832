"""The C# drop-in binding (compressiontests_amd/csharp/*.cs) against the C ABI it calls.

No dotnet exists here, so the binding is checked as text: every [StructLayout(Sequential)]
struct of Mgx.cs must have the C struct's fields in the same order with matching types and the
same byte offsets (C offsets from a gcc-compiled probe of include/mgx.h; C# offsets by the
sequential-layout rule, natural alignment), and every [DllImport] must name a function of the
header with the same number and kind of parameters.  The solver classes must expose the
reference's public API (OpenCLCsrGeometricMultigridSolver.cs:55-104,145,334,106;
OpenCLDuViGeometricMultigridSolver.cs:77-104; OpenCLCsrPcgSolver.cs:40-45,84,250,47) with the
OpenCL device and context arguments dropped, and the ISolver adapter must override what
PcgSolver2 overrides (PcgSolver2.cs:20-155) and keep its Factory."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CS = os.path.join(ROOT, "compressiontests_amd", "csharp")
HEADER = os.path.join(ROOT, "include", "mgx.h")

C_TO_CS = {"int32_t": "int", "int64_t": "long", "double": "double", "uint8_t": "byte", "float": "float"}
SIZE = {"int": 4, "long": 8, "double": 8, "byte": 1, "float": 4}


def strip_c_comments(t):
    return re.sub(r"/\*.*?\*/", "", t, flags=re.S)


def c_structs():
    text = strip_c_comments(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"(const\s+)?(\w+)\s+(.*)", decl)
            base, rest = m.group(2), m.group(3)
            for d in rest.split(","):
                d = d.strip()
                ptr = d.startswith("*")
                fields.append((d.lstrip("*").strip(), C_TO_CS.get(base, base) + ("*" if ptr else "")))
        out[name] = fields
    return out


def cs_structs(path):
    text = open(path).read()
    out = {}
    for name, body in re.findall(r"\[StructLayout\(LayoutKind\.Sequential\)\]\s*public\s+(?:unsafe\s+)?struct\s+(\w+)"
                                 r"\s*\{(.*?)\n\s*\}", text, flags=re.S):
        out[name] = re.findall(r"public\s+([\w*]+)\s+(\w+)\s*;", body)
    return out


def cs_offsets(fields):
    off, res, align = 0, [], 1
    for typ, name in fields:
        sz = 8 if typ.endswith("*") else SIZE[typ]
        off = (off + sz - 1) // sz * sz
        res.append((name, off))
        off += sz
        align = max(align, sz)
    return res, (off + align - 1) // align * align


@pytest.fixture(scope="module")
def c_offsets(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler to probe the struct layouts")
    structs = c_structs()
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for s in ("mgx_options", "mgx_stats", "mgx_csr_host"):
        for f, _ in structs[s]:
            lines.append(f'  printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
        lines.append(f'  printf("{s} sizeof %zu\\n", sizeof({s}));')
    lines.append("  return 0; }")
    d = tmp_path_factory.mktemp("probe")
    src, exe = d / "probe.c", d / "probe"
    src.write_text("\n".join(lines))
    subprocess.run([cc, "-std=c11", "-o", str(exe), str(src)], check=True)
    out = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, f, v = line.split()
        out.setdefault(s, {})[f] = int(v)
    return out


@pytest.mark.parametrize("name", ["mgx_options", "mgx_stats", "mgx_csr_host"])
def test_struct_layouts_match_the_header(name, c_offsets):
    c = c_structs()[name]
    cs = cs_structs(os.path.join(CS, "Mgx.cs"))[name]
    assert [f for f, _ in c] == [n for _, n in cs], name
    assert [t for _, t in c] == [t for t, _ in cs], name
    offs, size = cs_offsets(cs)
    assert {n: o for n, o in offs} == {f: c_offsets[name][f] for f, _ in c}
    assert size == c_offsets[name]["sizeof"]


def c_prototypes():
    text = strip_c_comments(open(HEADER).read())
    protos = {}
    for ret, name, args in re.findall(r"\b(int32_t|int|void|const char\s*\*)\s*(mgx_\w+)\s*\(([^)]*)\)\s*;", text):
        args = [a.strip() for a in args.split(",") if a.strip() and a.strip() != "void"]
        protos[name] = args
    return protos


def test_dllimports_match_prototypes():
    text = open(os.path.join(CS, "Mgx.cs")).read()
    protos = c_prototypes()
    imports = re.findall(r"\[DllImport\(Lib\)\]\s*public static extern\s+(?:unsafe\s+)?[\w*]+\s+(mgx_\w+)\s*\(([^)]*)\)",
                         text, flags=re.S)
    assert len(imports) >= 10
    for name, args in imports:
        assert name in protos, name
        cs_args = [a.strip() for a in args.split(",") if a.strip()]
        c_args = protos[name]
        assert len(cs_args) == len(c_args), name
        for ca, sa in zip(c_args, cs_args):
            c_ptr = "*" in ca or re.match(r"(mgx_solver|mgx_matrix|mgx_nodeblock|mgx_nodetransfer)\b", ca)
            s_ptr = "*" in sa or sa.startswith(("ref ", "out ", "IntPtr"))
            assert bool(c_ptr) == bool(s_ptr), (name, ca, sa)
            if not c_ptr:
                ctype = re.sub(r"\s+\w+$", "", ca).replace("const ", "").strip()
                assert sa.split()[0] == C_TO_CS[ctype], (name, ca, sa)


# the reference's public API of each solver (parameter names), without the OpenCL Device / context
REFERENCE_API = {
    "HipCsrGeometricMultigridSolver.cs": {
        "CreateSimpleV": ["model", "GaussSeidel", "maxCircleIterations", "coarseRelaxation", "convergenceTolerance",
                          "levelIterations"],
        "CreateDeepV": ["model", "GaussSeidel", "maxCircleIterations", "coarseRelaxation", "convergenceTolerance",
                        "depth", "levelIterations"],
        "HipCsrGeometricMultigridSolver": ["maxCircleIterations", "convergenceTolerance", "levelDown", "levelIterations",
                                           "GaussSeidel"],
        "Initialize": ["model", "coarseRelaxation"],
        "Solve": ["xInitialGuess"],
        "ReleaseOpenCLResources": [],
    },
    "HipDuViGeometricMultigridSolver.cs": {
        "CreateSimpleV": ["model", "GaussSeidel", "maxCircleIterations", "coarseRelaxation", "convergenceTolerance",
                          "levelIterations"],
        "CreateDeepV": ["model", "GaussSeidel", "maxCircleIterations", "coarseRelaxation", "convergenceTolerance",
                        "depth", "levelIterations"],
    },
    "HipCsrPcgSolver.cs": {
        "HipCsrPcgSolver": ["maxIterations", "convergenceTolerance"],
        "Initialize": ["model"],
        "Solve": ["xInitialGuess"],
        "ReleaseOpenCLResources": [],
    },
}


@pytest.mark.parametrize("fname", sorted(REFERENCE_API))
def test_solver_classes_expose_the_reference_api(fname):
    text = open(os.path.join(CS, fname)).read()
    for method, params in REFERENCE_API[fname].items():
        m = re.search(r"public\s+(?:new\s+)?(?:static\s+)?(?:unsafe\s+)?(?:[\w<>\[\](),?\s]+\s)?" + method + r"\s*\(([^)]*)\)",
                      text)
        assert m, (fname, method)
        names = [re.sub(r"\s*=.*", "", a).split()[-1] for a in m.group(1).split(",") if a.strip()]
        assert names == params, (fname, method, names)
    assert "Mgx.Check(" in text or fname == "HipDuViGeometricMultigridSolver.cs"


def test_solve_returns_the_reference_tuple():
    t = open(os.path.join(CS, "HipCsrGeometricMultigridSolver.cs")).read()
    assert re.search(r"public unsafe \(Vector, IterativeStatistics, double\[\]\) Solve\(Vector\? xInitialGuess\)", t)
    assert ": IOpenCLGeometricMultigridSolver" in t
    p = open(os.path.join(CS, "HipCsrPcgSolver.cs")).read()
    assert re.search(r"public unsafe \(Vector, IterativeStatistics\) Solve\(Vector\? xInitialGuess\)", p)


def test_pcg_solver2_adapter_keeps_the_reference_overrides():
    """HipPcgSolver2 derives from the base class PcgSolver2 derives from (with the DOK matrix type),
    overrides the same five members with the same signatures, throws the reference's exception on
    non-convergence, keeps the nested Factory's members, and drives the library through Mgx."""
    t = open(os.path.join(CS, "HipPcgSolver2.cs")).read()
    assert re.search(r"public class HipPcgSolver2 : SingleSubdomainSolverBase<DokRowMajor>", t)
    assert "namespace MGroup.Solvers.Iterative" in t
    for sig in (r"public override void HandleMatrixWillBeSet\(\)", r"public override void Initialize\(\)",
                r"public override void PreventFromOverwrittingSystemMatrices\(\)",
                r"public override (?:unsafe )?void Solve\(\)",
                r"protected override Matrix InverseSystemMatrixTimesOtherMatrix\(IReadOnlyMatrix otherMatrix\)"):
        assert re.search(sig, t), sig
    assert "throw new IterativeSolverNotConvergedException(" in t
    f = t[t.index("public class Factory"):]
    for member in ("public IDofOrderer DofOrderer", "public bool MatrixPatternWillNotBeModified",
                   "public HipPcgSolver2 BuildSolver(GlobalAlgebraicModel<DokRowMajor> model)",
                   "public GlobalAlgebraicModel<DokRowMajor> BuildAlgebraicModel(IModel model)"):
        assert member in f, member
    assert "new DokRowMajorMatrixAssembler(true)" in f
    for call in ("Mgx.mgx_solver_create(", "Mgx.mgx_solver_set_rhs(", "Mgx.mgx_solver_initialize(",
                 "Mgx.mgx_solver_solve(", "Mgx.mgx_solver_release(", "Mgx.SetLevel("):
        assert call in t, call
    assert "o.method = Mgx.METHOD_PCG_JACOBI;" in t and "o.stop = Mgx.STOP_REL_L2;" in t
