"""CPU checks of the reference-signature shims (icep_kapsel_amd/shims): the
library builds and exports every shim, the Particle layout seen from C++ is
the one the tests bind (variable.h:27-77, member for member), and the AoS <->
SoA marshalling round-trips every particle field the library keeps while
leaving the other members alone.  No device is needed."""
import ctypes
import os
import subprocess

import numpy as np

from icep_kapsel_amd import abi
from tests import shim_binding as sb

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_particle_layout_matches_header():
    members = [m for m, _ in sb._PARTICLE_MEMBERS]
    src = "#include <cstdio>\n#include <cstddef>\n#include \"reference_shims.hpp\"\nint main(){\n"
    src += 'printf("%zu\\n", sizeof(Particle));\n'
    for m in members:
        src += f'printf("%zu\\n", offsetof(Particle, {m}));\n'
    src += "return 0;}\n"
    tmp = "/tmp/kapsel_particle_layout"
    with open(tmp + ".cpp", "w") as f:
        f.write(src)
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "icep_kapsel_amd", "shims"), tmp + ".cpp", "-o",
                    tmp], check=True)
    vals = [int(v) for v in subprocess.run([tmp], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == ctypes.sizeof(sb.Particle)
    assert vals[1:] == [getattr(sb.Particle, m).offset for m in members]


def test_pack_unpack_round_trip():
    lib, fns = sb.load()
    n = 7
    rng = np.random.default_rng(4)
    arr = (sb.Particle * n)()
    for p in arr:
        for m, t in sb._PARTICLE_MEMBERS:
            if t is sb._V:
                getattr(p, m)[:] = list(rng.standard_normal(3))
        p.mass = 3.5
        p.spec = 2
    state = np.zeros((abi.P_NFIELDS, 3, n))
    fns["kapsel_shims_pack"](arr, n, sb.dptr(state))
    for f, m in sb.MEMBER_OF.items():
        assert np.array_equal(state[abi.PFIELD[f]].T, sb.member_array(arr, m))
    back = (sb.Particle * n)()
    for p in back:
        p.v_slip[:] = [9.0, 9.0, 9.0]
    state2 = state * 2.0
    fns["kapsel_shims_unpack"](sb.dptr(state2), n, back)
    for f, m in sb.MEMBER_OF.items():
        assert np.array_equal(sb.member_array(back, m), state2[abi.PFIELD[f]].T)
    assert np.all(sb.member_array(back, "v_slip") == 9.0)  # members the library does not keep stay as they were
