"""enginit's energy-coordinate mesh in the package (erdst_gpu/mesh.py)
against the oracle's restatement (bit for bit), EcdMesh and EcdInfo input,
the carry-over of &hist values between species, and the driver's checks of
what the GPU path supports."""
import numpy as np
import pytest

import oracle as O
from erdst_gpu import erdst
from erdst_gpu import inputs as I
from erdst_gpu import mesh as M

HIST = dict(ecdmin=-80.0, ecfmns=-0.20, ecdcen=0.0, eccore=20.0, ecdmax=1.0e11, eclbin=5.0e-2, ecfbin=2.0e-3,
            ec0bin=2.0e-4, finfac=10.0, pecore=200)


def hist_text(h, extra=""):
    body = ",\n".join(f"  {k} = {v!r}" for k, v in h.items())
    return f"&ene_param\n  slttype = 3\n/\n&hist\n{body}{extra}\n/\n"


@pytest.mark.parametrize("h", [HIST, dict(HIST, ecdmin=-20.0, eccore=5.0, ecdmax=1e4, pecore=100),
                               dict(HIST, ec0bin=1e-3, finfac=5.0, pecore=0)])
def test_mesh_equals_oracle(h):
    """Same counts; points equal to 2 ulp: the oracle is built for a host ISA
    with FMA and its compiler contracts base + factor*k into one rounding,
    which the Fortran statement (engproc.F90:222-228) allows either way."""
    mesh, pemax, pesoft = M.build_mesh(dict(h))
    ref, ref_soft = O.build_mesh(**h)
    assert (pemax, pesoft) == (ref.size, ref_soft)
    # a point is start + width*k: one rounding of the product (magnitude up
    # to |ecdmin|, ulp 1.4e-14 at 80) separates the two forms, nine orders
    # below the narrowest bin (ec0bin = 2e-4)
    np.testing.assert_allclose(mesh, ref, rtol=1e-14, atol=1e-13)


def test_species_meshes_from_hist(tmp_path):
    (tmp_path / "parameters_er").write_text(hist_text(HIST))
    uvmax, uvsoft, uvcrd, escrd = M.species_meshes(tmp_path / "parameters_er", 2)
    ref, soft = O.build_mesh(**HIST)
    assert list(uvmax) == [ref.size, ref.size] and list(uvsoft) == [soft, soft]
    np.testing.assert_allclose(uvcrd, np.concatenate([ref, ref]), rtol=1e-14, atol=1e-13)
    np.testing.assert_allclose(escrd, ref[:soft], rtol=1e-14, atol=1e-13)   # species 0 keeps its soft part


def test_hist_values_carry_over_between_species(tmp_path):
    """A variable the namelist leaves out keeps the value the previous
    species' mesh build shifted it to (engproc.F90:127-134, 170-176)."""
    h = dict(HIST)
    del h["ecdmin"]
    (tmp_path / "parameters_er").write_text(hist_text(h))
    with pytest.raises(I.InputError):
        M.species_meshes(tmp_path / "parameters_er", 1)        # no ecdmin at all
    state = dict(HIST)
    M.build_mesh(state)
    assert state["ecdmin"] == -80.0 - 10.0 * 2.0e-3            # shifted in place


def test_ecdmesh_block_replaces_species_mesh(tmp_path):
    (tmp_path / "parameters_er").write_text(hist_text(dict(HIST, pecore=200), ",\n  meshread = 1"))
    own = np.concatenate([np.linspace(-10, 5, 301), np.geomspace(5.5, 1e4, 50)])
    lines = ["0 4", "0 1 -1.0", "0 2 0.0", "0 3 1.0", "0 4 2.0", f"1 {own.size} 50"]
    lines += [f"1 {i + 1} {float(v)!r}" for i, v in enumerate(own)]
    (tmp_path / "EcdMesh").write_text("\n".join(lines) + "\n")
    uvmax, uvsoft, uvcrd, escrd = M.species_meshes(tmp_path / "parameters_er", 1,
                                                   ecdmesh_path=tmp_path / "EcdMesh")
    assert (uvmax[0], uvsoft[0]) == (own.size, own.size - 50)
    np.testing.assert_array_equal(uvcrd, own)
    np.testing.assert_array_equal(escrd, [-1.0, 0.0, 1.0, 2.0])
    # a non-increasing block halts eng_emf
    bad = lines[:6] + [f"1 {i + 1} {float(v)!r}" for i, v in enumerate(own[::-1])]
    (tmp_path / "EcdMesh").write_text("\n".join(bad) + "\n")
    with pytest.raises(I.InputError, match="eng_emf"):
        M.species_meshes(tmp_path / "parameters_er", 1, ecdmesh_path=tmp_path / "EcdMesh")


def test_ecdinfo_overrides_hist(tmp_path):
    (tmp_path / "parameters_er").write_text(hist_text(HIST, ",\n  ecprread = 1"))
    (tmp_path / "EcdInfo").write_text("# species eclbin ecfbin ec0bin finfac ecdmin ecfmns ecdcen eccore ecdmax pecore\n"
                                      "1 0.1 0.004 0.0004 10.0 -40.0 -0.4 0.0 10.0 1e6 100\n")
    uvmax, uvsoft, uvcrd, _ = M.species_meshes(tmp_path / "parameters_er", 1, ecdinfo_path=tmp_path / "EcdInfo")
    ref, soft = O.build_mesh(ecdmin=-40.0, ecfmns=-0.4, ecdcen=0.0, eccore=10.0, ecdmax=1e6, eclbin=0.1,
                             ecfbin=0.004, ec0bin=0.0004, finfac=10.0, pecore=100)
    np.testing.assert_allclose(uvcrd, ref, rtol=1e-14, atol=1e-13)
    assert uvsoft[0] == soft


@pytest.mark.parametrize("over,msg", [(dict(cltype=1), "Ewald"), (dict(boxshp=2), "boxshp"), (dict(insstructure=1), "structure"),
                                      (dict(insposition=0), "as they are")])
def test_driver_rejects_what_the_gpu_path_does_not_run(over, msg):
    p = dict(slttype=3, cltype=2, boxshp=1, insposition=1, insorient=1, insstructure=0)
    p.update(over)
    with pytest.raises(I.InputError, match=msg):
        erdst._check_supported(p)


@pytest.mark.parametrize("pecore,ok", [(200, True), (0, False)])
def test_ecdmesh_species_without_a_block(tmp_path, pecore, ok):
    """With meshread, a species with no EcdMesh block keeps its &hist mesh
    when that mesh has a core region: the line counter is the build loop's
    index, which then ends at pemax (engproc.F90:222-228, 238-261); without
    a core region the count check fails (eng_emf)."""
    (tmp_path / "parameters_er").write_text(hist_text(dict(HIST, pecore=pecore), ",\n  meshread = 1"))
    own = np.linspace(-10, 5, 301)
    lines = [f"1 {own.size} 0"] + [f"1 {i + 1} {float(v)!r}" for i, v in enumerate(own)]
    (tmp_path / "EcdMesh").write_text("\n".join(lines) + "\n")
    if ok:
        _, _, uvcrd, escrd = M.species_meshes(tmp_path / "parameters_er", 1, ecdmesh_path=tmp_path / "EcdMesh")
        ref, soft = O.build_mesh(**dict(HIST, pecore=pecore))
        np.testing.assert_allclose(escrd, ref[:soft], rtol=1e-14, atol=1e-13)
        np.testing.assert_array_equal(uvcrd, own)
    else:
        with pytest.raises(I.InputError, match="eng_emf"):
            M.species_meshes(tmp_path / "parameters_er", 1, ecdmesh_path=tmp_path / "EcdMesh")


def test_region_distance_sphere_and_slab():
    """check_mol_configuration's distances (engproc.F90:1228-1249): the solute
    centre of mass relative to the hostspec aggregate's, as a norm (sphere),
    as the component along c (slab) and its magnitude (symmetric slab)."""
    from types import SimpleNamespace
    topo = SimpleNamespace(mol_begin=np.array([0, 2, 4, 6]), nmol=3, moltype=np.array([1, 2, 2]))
    mass = np.array([1.0, 3.0, 2.0, 2.0, 1.0, 1.0], np.float32)
    xyz = np.array([[0, 0, 0], [4, 0, 0], [1, 1, 1], [3, 1, 1], [0, 0, 5], [0, 0, 7]], np.float32)
    sysin = SimpleNamespace(topology=topo, mass=mass)
    slt = np.array([3.0, 0.0, 0.0])                     # (0*1 + 4*3) / 4
    agg = (2 * np.array([1, 1, 1]) + 2 * np.array([3, 1, 1]) + np.array([0, 0, 5]) + np.array([0, 0, 7])) / 6
    dx = slt - agg
    cell = np.diag([20.0, 20.0, 20.0])
    p = dict(insposition=2, hostspec={1: 2})
    assert erdst.region_distance(p, sysin, xyz, cell, 0) == pytest.approx(np.linalg.norm(dx), rel=1e-12)
    assert erdst.region_distance(dict(p, insposition=3), sysin, xyz, cell, 0) == pytest.approx(dx[2], rel=1e-12)
    assert erdst.region_distance(dict(p, insposition=4), sysin, xyz, cell, 0) == pytest.approx(abs(dx[2]))
    sheared = np.array([[20.0, 0.0, 5.0], [0.0, 20.0, 0.0], [0.0, 0.0, 20.0]])   # c = (5, 0, 20)
    s3 = np.linalg.solve(sheared, dx)[2]
    assert erdst.region_distance(dict(p, insposition=3), sysin, xyz, sheared, 0) == pytest.approx(
        s3 * np.hypot(5.0, 20.0), rel=1e-12)
