"""The public headers are a C API too: each compiles on its own as C99 (gcc,
HIP's platform flags) and as C++, so cgo / plain C hosts can include them;
the C host program built from them exists."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
HEADERS = sorted(f for f in os.listdir(INC) if f.endswith(".h"))
FLAGS = ["-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", INC, "-fsyntax-only", "-Wall", "-Wextra",
         "-Werror"]


@pytest.mark.parametrize("lang,std", [("c", "-std=c99"), ("c", "-std=c11"), ("c++", "-std=c++17")])
@pytest.mark.parametrize("header", HEADERS)
def test_header_compiles_alone(tmp_path, header, lang, std):
    src = tmp_path / ("t.c" if lang == "c" else "t.cpp")
    src.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
    cc = "gcc" if lang == "c" else "g++"
    r = subprocess.run([cc, "-x", lang, std] + FLAGS + [str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_every_header_together_in_c(tmp_path):
    src = tmp_path / "all.c"
    src.write_text("".join(f'#include "{h}"\n' for h in HEADERS) +
                   "int main(void) { Parameters *p = gasal2_parameters_new(); gasal2_parameters_free(p);\n"
                   "  comp_start s = WITH_TB; algo_type a = LOCAL; data_source d = TARGET; Bool b = TRUE;\n"
                   "  return (int)s + (int)a + (int)d + (int)b; }\n")
    r = subprocess.run(["gcc", "-x", "c", "-std=c99"] + FLAGS + [str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_c_host_built():
    assert os.access(os.path.join(ROOT, "gasal2_amd", "c_host"), os.X_OK), \
        "gasal2_amd/c_host missing: run python -m gasal2_amd.build"
