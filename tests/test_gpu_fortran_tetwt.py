"""The tetrahedron weights called from Fortran through ecalj_gpu_mod.f90
(tests/fortran/tetwt_call.f90, built by build()): ecalj_tetwt_full once for
the sizes and once with arrays of those sizes, as gettetwt's caller would,
with npm = 2, core states and tetrahedron_matrix_linear on and off, against
the oracle.  Lists bit-exact; weights within 1e-8 of the largest (the bar of
the C-ABI tests of the same entry point)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from ecalj_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "tetwt_call")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran caller not built (build() needs flang)")
@pytest.mark.parametrize("npm,ncore,ml", [(1, 0, 0), (2, 2, 0), (2, 2, 1)])
def test_fortran_tetwt(orc, tmp_path, npm, ncore, ml):
    print(f"fortran tetwt: npm {npm} core {ncore} matrix_linear {ml}", flush=True)
    sys.stdout.flush()
    m = synth.tetra_mesh(3, 2, 2)
    ek1, ek2, ef = synth.tetra_bands(m, 5, 9, seed=31)
    ecore = np.array([-32.0, -19.5][:ncore]) / synth.RYDBERG_EV
    full = ek1 if ncore == 0 else np.vstack([np.repeat(ecore[:, None], ek1.shape[1], 1), ek1])
    fr = synth.tetra_frhis(full, ek2)
    nband, nqbz = ek1.shape
    idt, qw = np.asfortranarray(m.idtetf, np.int32), np.asfortranarray(m.qbzw, np.float64)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([nband, nqbz, idt.shape[1], qw.shape[1], len(fr) - 1, npm, ncore, ml], np.int32).tofile(f)
        np.array([ef, 1e10], np.float64).tofile(f)
        for a in (idt, qw, np.asarray(m.ib1bz, np.int32), np.asfortranarray(ek1, np.float64),
                  np.asfortranarray(ek2, np.float64), np.asfortranarray(m.qlat, np.float64), np.asarray(fr)):
            np.asarray(a).ravel(order="F").tofile(f)
        if ncore:
            np.asarray(ecore, np.float64).tofile(f)
    r = subprocess.run([EXE, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    o = orc.tetwt(m, ek1, ek2, ef, fr, npm=npm, ecore=ecore if ncore else None, matrix_linear=bool(ml))
    with open(fout, "rb") as f:
        nbnbx = int(np.fromfile(f, np.int32, 1)[0])
        nhwtot = int(np.fromfile(f, np.int64, 1)[0])
        nbnb = np.fromfile(f, np.int32, nqbz * npm).reshape((nqbz, npm), order="F")
        lists = [np.fromfile(f, np.int32, nbnbx * nqbz * npm).reshape((nbnbx, nqbz, npm), order="F")
                 for _ in range(5)]
        whw = np.fromfile(f, np.float64, nhwtot)
    assert nbnbx == o["nbnbx"] and nhwtot == o["nhwtot"]
    if npm == 1:
        nbnb, lists = nbnb[:, 0], [a[:, :, 0] for a in lists]
    np.testing.assert_array_equal(nbnb, o["nbnb"])
    for key, a in zip(("n1b", "n2b", "ihw", "nhw", "jhw"), lists):
        np.testing.assert_array_equal(a, o[key], err_msg=key)
    assert np.abs(whw - o["whw"]).max() <= 1e-8 * np.abs(o["whw"]).max()
    assert abs(whw.sum() - o["whw"].sum()) <= 1e-12 * abs(o["whw"].sum())
