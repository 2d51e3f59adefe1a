"""The ctypes mirror of include/vx_solver.h has the C layout: every struct's
size and every field's offset, as gcc lays them out from the header, equal
the Python mirror's."""
import ctypes
import os
import shutil
import subprocess

import pytest

from vxsim import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = {
    "vx_scene_desc": _abi.SceneDesc,
    "vx_solve_params": _abi.SolveParams,
    "vx_solve_report": _abi.SolveReport,
    "vx_fracture_params": _abi.FractureParams,
    "vx_fracture_report": _abi.FractureReport,
    "vx_step_params": _abi.StepParams,
    "vx_step_report": _abi.StepReport,
    "vx_batch_scene_report": _abi.BatchSceneReport,
    "vx_contact": _abi.Contact,
    "vx_contact_report": _abi.ContactReport,
    "vx_kernel_stats": _abi.KernelStats,
    "vx_kernel_times": _abi.KernelTimes,
    "vx_tool_desc": _abi.ToolDesc,
    "vx_tool_state": _abi.ToolState,
}


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    d = tmp_path_factory.mktemp("layout")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vx_solver.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines += ["return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines))
    exe = d / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    return dict((k, int(v)) for k, v in (line.split() for line in out.splitlines()))


@pytest.mark.parametrize("cname", sorted(STRUCTS))
def test_struct_layout_matches_header(c_layout, cname):
    py = STRUCTS[cname]
    assert ctypes.sizeof(py) == c_layout[cname]
    for f in py._fields_:
        assert getattr(py, f[0]).offset == c_layout["%s.%s" % (cname, f[0])], f[0]


def test_header_fields_are_all_mirrored():
    """Every field the header declares for these structs is in the mirror
    (gcc rejects offsetof of a field the header lacks, the converse is
    checked here by counting the header's members)."""
    import re
    text = open(os.path.join(ROOT, "include", "vx_solver.h")).read()
    for cname, py in STRUCTS.items():
        m = re.search(r"typedef struct %s \{(.*?)\} %s;" % (cname, cname), text, re.S)
        assert m, cname
        body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
        names = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            # "double a, b;" and "const float *p;" and "double gravity[3];"
            parts = [decl.split()[-1]] if "," not in decl else [p.strip() for p in decl.split(",")]
            for p in parts:
                names.append(re.sub(r"[\*\[\]0-9]", "", p.split()[-1]))
        assert names == [f[0] for f in py._fields_], cname
