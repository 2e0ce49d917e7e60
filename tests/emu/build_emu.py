"""CPU build of the product's HIP sources (test infrastructure).

Compiles stress-particle-sph_amd/csrc/sps_api.hip (which includes the
kernels) with g++ against the host stand-in in tests/emu/hip/, after three
textual rewrites of a copy of the sources:

* each launch `kern<<<grid, block, lds, stream>>>(args)` becomes
  `emu::launch_k(grid, block, lds, sync, kern, args)`, which evaluates the
  arguments once, as a launch does;
* the dynamic LDS array becomes the launch's buffer of exactly the requested
  size, so an access past it is reported;
* a kernel's local layout reference used as a template argument is replaced
  by the layout it names (g++ does not take a local reference there).

The library implements include/sps.h like the product and the oracle, so the
Python package drives it the same way.  Built with AddressSanitizer by
default: run the Python process with LD_PRELOAD of libasan.

    python tests/emu/build_emu.py [out.so] [--no-asan]
"""
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(ROOT, "stress-particle-sph_amd", "csrc")
BUILD = os.path.join(HERE, "_build")

# kernels that use barriers or wave exchanges (their threads run as fibers)
SYNC = ("k_bounds_partial", "k_bounds_final", "k_count", "k_spu_nodes", "k_sp_update", "k_deriv_nodes", "k_permute",
        "k_scan_local", "k_scan_sums", "k_max_disp", "k_gemm_bf16", "k_mlp_loss", "k_mlp_colsum",
        "k_surrogate_apply", "k_xsph", "k_av_nodes", "k_cell_count", "k_cell_scatter")


def split_top(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        depth += ch in "([{"
        depth -= ch in ")]}"
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    out.append(cur)
    return [c.strip() for c in out]


def rewrite_launches(s):
    out, i = [], 0
    while True:
        k = s.find("<<<", i)
        if k < 0:
            out.append(s[i:])
            return "".join(out)
        m = re.search(r"([A-Za-z_][A-Za-z_0-9:]*(?:<[^<>]*>)?)\s*$", s[:k])
        name, start = m.group(1), m.start(1)
        e = s.find(">>>", k)
        cfg = split_top(s[k + 3:e])
        p = s.find("(", e)
        depth, q = 0, p
        while True:
            depth += s[q] == "("
            depth -= s[q] == ")"
            if depth == 0:
                break
            q += 1
        args = s[p + 1:q]
        if name == "KERN":          # the launch macro: decide by the kernel's name
            sync = "emu::needs_sync(#KERN)"
        else:
            sync = "true" if name.split("<")[0].startswith(SYNC) else "false"
        out.append(s[i:start])
        out.append(f"emu::launch_k({cfg[0]}, {cfg[1]}, {cfg[2]}, {sync}, {name}, {args})")
        i = q + 1


def generate(gen):
    os.makedirs(gen, exist_ok=True)
    api = open(os.path.join(SRC, "sps_api.hip")).read()
    api = api.replace('#include "../../include/sps.h"', '#include "sps.h"')
    api = '#include "hip/hip_runtime.h"\n#define sps_lds (::emu::lds_base)\n' + rewrite_launches(api)
    open(os.path.join(gen, "sps_api_emu.cpp"), "w").write(api)
    host = rewrite_launches(open(os.path.join(SRC, "sps_surrogate_host.hip")).read())
    open(os.path.join(gen, "sps_surrogate_host.hip"), "w").write(host)
    decl = "extern __shared__ __align__(16) unsigned char sps_lds[];"
    for name in ("sps_kernels.hip", "sps_mixed.hip", "sps_surrogate.hip"):
        ker = open(os.path.join(SRC, name)).read()
        if name == "sps_kernels.hip" and decl not in ker:
            raise RuntimeError("dynamic LDS declaration not found")
        lines, cur = [], None
        for ln in ker.replace(decl, "").split("\n"):
            m = re.search(r"constexpr const LdsLayout &LY = (LAY_\w+);", ln)
            if m:
                cur = m.group(1)
            elif cur:
                ln = ln.replace("<LY>", "<" + cur + ">")
            lines.append(ln)
        text = "\n".join(lines)
        # g++ spells the device vector types with vector_size
        text = re.sub(r"typedef (short|float) (\w+) __attribute__\(\(ext_vector_type\(4\)\)\);",
                      r"typedef \1 \2 __attribute__((vector_size(4 * sizeof(\1))));", text)
        open(os.path.join(gen, name), "w").write(text)
    for name in ("sps_device.h", "sps_tiles.h"):
        open(os.path.join(gen, name), "w").write(open(os.path.join(SRC, name)).read())


def build(out=None, asan=True, defines=()):
    """The emulated library; `defines` are extra -D definitions (e.g. another
    tile width), each build with its own generated sources."""
    tag = "_".join(d.replace("=", "") for d in defines)
    gen = os.path.join(BUILD, "gen" + ("_" + tag if tag else ""))
    generate(gen)
    out = out or os.path.join(BUILD, f"libsps_emu{'_' + tag if tag else ''}.so")
    flags = ["-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
             "-U_FORTIFY_SOURCE", "-D_FORTIFY_SOURCE=0", "-Wno-unknown-pragmas", "-Wno-unused-result",
             "-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-I" + gen]
    if asan:
        flags += ["-fsanitize=address", "-fno-omit-frame-pointer"]
    flags += ["-D" + d for d in defines]
    subprocess.run(["g++"] + flags + ["-o", out, os.path.join(gen, "sps_api_emu.cpp"),
                                      os.path.join(HERE, "emu_rt.cpp")], check=True)
    return out


if __name__ == "__main__":
    args = [a for a in sys.argv[1:] if not a.startswith("--")]
    print(build(args[0] if args else None, asan="--no-asan" not in sys.argv))
