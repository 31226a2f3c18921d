"""Drop-in host layer (magritte_amd.core: Io, Model, Simulation).

CPU: a model written with IoText and read back into a fresh Model carries the
same members (the io layout of src/Model/model.cpp:54-86 and the component
read()/write() functions); the SetOnce parameters raise as the reference's
(src/Tools/setOnce.hpp:15-85); a Simulation without a spectral discretisation
returns -1 from compute_level_populations (sim_lines.tpp:66-71) without
touching a GPU.

GPU: the Simulation methods, called in the order of the reference's own
tests (compute_spectral_discretisation, compute_boundary_intensities,
compute_LTE_level_populations, compute_radiation_field,
compute_level_populations), reproduce the oracle's J within 1e-5 and its
converged populations within 1e-5.
"""
import os

import numpy as np
import pytest

from magritte_amd.setup import synthetic

core = pytest.importorskip("magritte_amd.core", reason="host layer not built")
from magritte_amd.host import to_simulation  # noqa: E402


def _small():
    return synthetic.envelope_model(n=6, seed=3)


def test_set_once_semantics():
    p = core.Parameters()
    with pytest.raises(RuntimeError, match="GetBeforeSet"):
        p.ncells()
    p.set_ncells(10)
    p.set_ncells(10)
    with pytest.raises(RuntimeError, match="DoubleSet"):
        p.set_ncells(11)
    assert p.ncells() == 10


def _io(kind, tmp_path):
    if kind == "text":
        return core.IoText(str(tmp_path) + "/model/")
    if not hasattr(core, "IoPython"):
        pytest.skip("HDF5 C library not found at build time")
    return core.IoPython("hdf5", str(tmp_path) + "/model.hdf5")


@pytest.mark.parametrize("kind", ["text", "hdf5"])
def test_io_round_trip(tmp_path, kind):
    """Model::write then Model::read through IoText and through the HDF5
    back end (IoPython("hdf5"), io_python_hdf5.py) reproduce every member."""
    m = _small()
    sim = to_simulation(m)
    io = _io(kind, tmp_path)
    sim.write(io)
    back = core.Model()
    back.read(io)
    assert back.parameters.ncells() == m.ncells
    assert back.parameters.nrays() == m.nrays
    assert back.parameters.nboundary() == m.nboundary
    np.testing.assert_array_equal(np.array(back.geometry.cells.position), m.position)
    np.testing.assert_array_equal(np.array(back.geometry.cells.velocity), m.velocity)
    nb = np.concatenate([np.asarray(x, dtype=np.int64) for x in back.geometry.cells.neighbors])
    np.testing.assert_array_equal(nb, m.neighbors)
    np.testing.assert_array_equal(np.array(back.geometry.cells.n_neighbors), m.n_neighbors)
    np.testing.assert_array_equal(np.array(back.geometry.rays.rays), m.rays)
    np.testing.assert_array_equal(np.array(back.geometry.rays.weights), m.weights)
    np.testing.assert_array_equal(np.array(back.geometry.boundary.boundary2cell_nr), m.boundary2cell_nr)
    np.testing.assert_array_equal(np.array(back.thermodynamics.temperature.gas), m.temperature)
    np.testing.assert_array_equal(np.array(back.thermodynamics.turbulence.vturb2), m.vturb2)
    np.testing.assert_array_equal(np.array(back.chemistry.species.abundance), m.abundance)
    ls = back.lines.lineProducingSpecies[0]
    ld = m.linedata
    assert ls.linedata.nlev == ld.nlev and ls.linedata.nrad == ld.nrad
    np.testing.assert_array_equal(np.array(ls.linedata.energy), ld.energy)
    np.testing.assert_array_equal(np.array(ls.linedata.A), ld.A)
    np.testing.assert_array_equal(np.array(ls.linedata.irad), ld.irad)
    assert len(ls.linedata.colpar) == len(ld.colpar)
    for a, b in zip(ls.linedata.colpar, ld.colpar):
        np.testing.assert_array_equal(np.array(a.Cd), b.Cd)
        np.testing.assert_array_equal(np.array(a.Ce), b.Ce)
        np.testing.assert_array_equal(np.array(a.tmp), b.tmp)
        assert a.num_col_partner == b.num
    np.testing.assert_array_equal(np.array(ls.quadrature.roots), m.roots)
    np.testing.assert_array_equal(np.array(ls.quadrature.weights), m.qweights)
    # the antipodes are set up on read (rays.cpp:31-56)
    rays = np.array(back.geometry.rays.rays)
    for r, a in enumerate(back.geometry.rays.antipod):
        np.testing.assert_allclose(rays[a], -rays[r], atol=1e-12)


@pytest.mark.parametrize("kind", ["text", "hdf5"])
def test_io_round_trip_of_two_species(tmp_path, kind):
    """A model with Lines/LineProducingSpecies_0 and _1 reads back both
    species (Lines::read, lines.cpp:19-40) with their line data and quadratures,
    and the merged sorted line list."""
    m = synthetic.two_species_model(n=4, nside=1, second_nu=200.0e9)
    sim = to_simulation(m)
    io = _io(kind, tmp_path)
    sim.write(io)
    back = core.Model()
    back.read(io)
    assert back.parameters.nlspecs() == 2
    assert back.parameters.nlines() == m.nlines
    for ls, ld in zip(back.lines.lineProducingSpecies, m.species_linedata):
        assert ls.linedata.nlev == ld.nlev and ls.linedata.nrad == ld.nrad and ls.linedata.num == ld.num
        np.testing.assert_array_equal(np.array(ls.linedata.frequency), ld.frequency)
        np.testing.assert_array_equal(np.array(ls.linedata.energy), ld.energy)
    freqs = np.concatenate([ld.frequency for ld in m.species_linedata])
    np.testing.assert_array_equal(np.array(back.lines.line), np.sort(freqs))


@pytest.mark.gpu
def test_two_species_simulation_matches_oracle(tmp_path):
    """The drop-in Simulation with two line-producing species: J and the
    converged populations of each species against the oracle; error_mean and
    error_max get one entry per species and iteration, and every species
    writes its populations (sim_lines.tpp:78-83, 130-151)."""
    from oracle.py_oracle import Oracle
    m = synthetic.two_species_model(n=5, nside=1, seed=4)
    print("\n[host] two species: Simulation vs oracle", flush=True)
    orc = Oracle(m).setup()
    sim = to_simulation(m)
    assert sim.compute_spectral_discretisation() == 0
    assert sim.compute_boundary_intensities() == 0
    assert sim.compute_LTE_level_populations() == 0
    assert sim.compute_radiation_field() == 0
    orc.compute_radiation_field()
    J_o = orc.J().ravel()
    assert float(np.max(np.abs(np.array(sim.radiation.J) - J_o) / np.abs(J_o))) < 1e-5
    for l, ls in enumerate(sim.lines.lineProducingSpecies):
        assert np.array(ls.nr_line).shape == (m.ncells, m.species_linedata[l].nrad, m.nquads)
    sim.convergence_criterion = core.CRIT_MAX
    it = sim.compute_level_populations(core.IoText(str(tmp_path) + "/"), True, 60)
    it_o = orc.compute_level_populations(True, 60, 1)
    assert 0 < it < 60 and abs(it - it_o) <= 1
    for ls, po in zip(sim.lines.lineProducingSpecies, orc.populations()):
        pg = np.array(ls.population)
        assert float(np.max(np.abs(pg - po.ravel()) / np.abs(po.ravel()))) < 1e-5
    assert len(sim.error_max) == 2 * it
    for l in (0, 1):
        written = os.listdir(str(tmp_path) + f"/Lines/LineProducingSpecies_{l}")
        assert f"population_rank_0_iteration_{it}.txt" in written


@pytest.mark.parametrize("kind", ["text", "hdf5"])
def test_io_round_trip_of_the_iteration_state(tmp_path, kind):
    """What a resumed ALI needs (LineProducingSpecies::read/write,
    lineProducingSpecies.cpp:21-134, 179-206): population, population_prev1..3,
    population_tot, J_lin and J_eff come back bit for bit, and the per-iteration
    dumps (write_populations with a tag, sim_lines.tpp:78-83, 147-151) read back
    through read_populations."""
    m = _small()
    sim = to_simulation(m)
    ls = sim.lines.lineProducingSpecies[0]
    V, K, N = m.linedata.nlev, m.nrad, m.ncells
    rng = np.random.default_rng(17)
    vals = {}
    for name in ("population", "population_prev1", "population_prev2", "population_prev3"):
        vals[name] = rng.lognormal(10.0, 3.0, N * V)
        setattr(ls, name, vals[name].tolist())
    vals["population_tot"] = rng.lognormal(10.0, 3.0, N)
    ls.population_tot = vals["population_tot"].tolist()
    vals["Jlin"] = rng.lognormal(-40.0, 2.0, (N, K))
    vals["Jeff"] = rng.lognormal(-40.0, 2.0, (N, K))
    ls.Jlin = vals["Jlin"].tolist()
    ls.Jeff = vals["Jeff"].tolist()
    io = _io(kind, tmp_path)
    sim.write(io)
    ls.write_populations(io, 0, "_rank_0_iteration_3")
    back = core.Model()
    back.read(io)
    lb = back.lines.lineProducingSpecies[0]
    for name in ("population", "population_prev1", "population_prev2", "population_prev3", "population_tot"):
        np.testing.assert_array_equal(np.array(getattr(lb, name)), vals[name], err_msg=name)
    np.testing.assert_array_equal(np.array(lb.Jlin), vals["Jlin"])
    np.testing.assert_array_equal(np.array(lb.Jeff), vals["Jeff"])
    lb.population = [0.0] * (N * V)
    assert lb.read_populations(io, 0, "_rank_0_iteration_3") == 0
    np.testing.assert_array_equal(np.array(lb.population), vals["population"])


def test_hdf5_layout(tmp_path):
    """The HDF5 file has the reference's layout: "<group>/.<name>" are
    attributes (root attributes for the parameters), lists and arrays are
    datasets at their paths (io_python_hdf5.py), booleans are words."""
    if not hasattr(core, "IoPython"):
        pytest.skip("HDF5 C library not found at build time")
    import shutil
    import subprocess
    h5ls = shutil.which("h5ls") or "/opt/conda/bin/h5ls"
    if not os.path.exists(h5ls):
        pytest.skip("h5ls not available")
    m = _small()
    sim = to_simulation(m)
    path = str(tmp_path) + "/model.hdf5"
    sim.write(core.IoPython("hdf5", path))
    listing = subprocess.run([h5ls, "-r", path], capture_output=True, text=True, check=True).stdout
    assert f"/Geometry/Cells/position Dataset {{{m.ncells}, 3}}" in listing
    assert f"/Geometry/Cells/neighbors Dataset {{{len(m.neighbors)}}}" in listing
    assert "/Lines/LineProducingSpecies_0/Linedata Group" in listing
    h5dump = os.path.join(os.path.dirname(h5ls), "h5dump")
    attrs = subprocess.run([h5dump, "-A", path], capture_output=True, text=True, check=True).stdout
    assert 'ATTRIBUTE "ncells"' in attrs and 'ATTRIBUTE "use_scattering"' in attrs
    assert 'GROUP "Linedata"' in attrs and 'ATTRIBUTE "num"' in attrs
    assert '"false"' in attrs   # booleans are words
    io = core.IoPython("hdf5", path)
    assert io.read_number(".ncells") == m.ncells
    assert io.implementation == "io_python_hdf5"
    with pytest.raises(Exception):
        core.IoPython("text", path)


def test_level_populations_without_spectral_discretisation_returns_minus_one(tmp_path):
    sim = to_simulation(_small())
    assert sim.compute_level_populations(core.IoText(str(tmp_path) + "/"), True, 10) == -1


def test_boundary_intensities_size_mismatch_returns_minus_one():
    sim = to_simulation(_small())
    assert sim.compute_boundary_intensities([1.0, 2.0]) == -1


@pytest.mark.gpu
def test_simulation_matches_oracle(tmp_path):
    from oracle.py_oracle import Oracle
    m = synthetic.envelope_model(n=8, seed=7)
    print("\n[host] envelope8 seed 7: Simulation vs oracle", flush=True)
    orc = Oracle(m).setup()
    sim = to_simulation(m)
    assert sim.compute_spectral_discretisation() == 0
    assert sim.compute_boundary_intensities() == 0
    assert sim.compute_LTE_level_populations() == 0
    nu_o, _ = orc.frequencies()
    np.testing.assert_array_equal(np.array(sim.radiation.frequencies.nu).ravel(), nu_o.ravel())
    assert sim.compute_radiation_field() == 0
    orc.compute_radiation_field()
    J_o = orc.J().ravel()
    J_g = np.array(sim.radiation.J)
    assert float(np.max(np.abs(J_g - J_o) / np.abs(J_o))) < 1e-5
    assert len(sim.timings) == 1 and sim.timings[0] > 0.0
    ls = sim.lines.lineProducingSpecies[0]
    Ls = np.array([[x[0] if len(x) else 0.0 for x in row] for row in ls.Lambda.Ls])
    assert Ls.shape == (m.ncells, m.nrad) and np.all(Ls > 0.0)

    sim.convergence_criterion = core.CRIT_MAX
    it = sim.compute_level_populations(core.IoText(str(tmp_path) + "/"), True, 60)
    it_o = orc.compute_level_populations(True, 60, 1)
    assert 0 < it < 60 and abs(it - it_o) <= 2
    pop_g = np.array(ls.population)
    pop_o = orc.populations().ravel()
    assert float(np.max(np.abs(pop_g - pop_o) / np.abs(pop_o))) < 1e-5
    assert len(sim.error_max) == it
    # per-iteration population files, tag "_rank_0_iteration_<k>" (sim_lines.tpp:149-151)
    written = os.listdir(str(tmp_path) + "/Lines/LineProducingSpecies_0")
    assert f"population_rank_0_iteration_{it}.txt" in written


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["text", "hdf5"])
def test_checkpoint_resume_equals_uninterrupted_run(tmp_path, kind):
    """An ALI stopped after five iterations (four normal ones and an Ng
    step, after which the reference's schedule starts again), written with
    Model::write, read into a fresh Simulation and continued, ends with the
    populations, history and total iteration count of an uninterrupted run
    (sim_lines.tpp:61-164; the state read back is lineProducingSpecies.cpp:21-134).
    The model needs 11 iterations, so the resumed part runs a second Ng step."""
    m = synthetic.envelope_model(n=10, n0=1.0e12, T0=1000.0, nside=1, seed=3, abundance=1.0e-3)
    print(f"\n[host] checkpoint/resume through {kind}", flush=True)

    def fresh():
        sim = to_simulation(m)
        assert sim.compute_spectral_discretisation() == 0
        assert sim.compute_boundary_intensities() == 0
        assert sim.compute_LTE_level_populations() == 0
        sim.convergence_criterion = core.CRIT_MAX
        return sim

    def state(sim):
        ls = sim.lines.lineProducingSpecies[0]
        return [np.array(getattr(ls, n)) for n in
                ("population", "population_prev1", "population_prev2", "population_prev3")]

    whole = fresh()
    it_whole = whole.compute_level_populations(core.IoText(str(tmp_path) + "/whole/"), True, 60)
    assert it_whole >= 10

    first = fresh()
    assert first.compute_level_populations(core.IoText(str(tmp_path) + "/first/"), True, 5) == 5
    (tmp_path / "checkpoint").mkdir()
    io = _io(kind, tmp_path / "checkpoint")
    first.write(io)
    del first

    resumed = core.Simulation()
    resumed.read(io)
    assert resumed.compute_spectral_discretisation() == 0
    assert resumed.compute_boundary_intensities() == 0
    resumed.convergence_criterion = core.CRIT_MAX
    it_rest = resumed.compute_level_populations(core.IoText(str(tmp_path) + "/rest/"), True, 60)
    print(f"[host] uninterrupted {it_whole} iterations, resumed 5 + {it_rest}", flush=True)
    assert 5 + it_rest == it_whole
    for a, b in zip(state(resumed), state(whole)):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(np.array(resumed.error_max), np.array(whole.error_max)[5:])


@pytest.mark.gpu
def test_lambda_lists_with_bands_match_oracle():
    """With n_off_diag = 3 the Simulation's Lambda lists hold, per (p, k), the
    element nr = p first and then the band elements in the order
    RayPair::update_Lambda first adds them (Lambda::add_element,
    lambda.tpp:130-143), as the oracle's lists; the band values are FP64."""
    from oracle.py_oracle import Oracle
    m = synthetic.envelope_model(n=5, nside=1, nquads=11, nlev=4, seed=3)
    m.n_off_diag = 3
    print("\n[host] Lambda lists with bands", flush=True)
    orc = Oracle(m).setup()
    orc.compute_radiation_field()
    sim = to_simulation(m)
    assert sim.compute_spectral_discretisation() == 0
    assert sim.compute_boundary_intensities() == 0
    assert sim.compute_LTE_level_populations() == 0
    assert sim.compute_radiation_field() == 0
    lam = sim.lines.lineProducingSpecies[0].Lambda
    band_o, diag_o = orc.lambda_band(), orc.lambda_diag()
    n_band = 0
    for p in range(m.ncells):
        for k in range(m.nrad):
            nr, Ls = list(lam.nr[p][k]), list(lam.Ls[p][k])
            assert nr[0] == p and nr[1:] == [c for c, _ in band_o[p][k]], (p, k)
            assert Ls[0] == pytest.approx(diag_o[p, k], rel=1e-5)
            for got, (_, want) in zip(Ls[1:], band_o[p][k]):
                assert got == pytest.approx(want, rel=1e-10)
            assert lam.size[k + m.nrad * p] == len(nr)
            n_band += len(nr) - 1
    assert n_band > 0


def _read_text_array(path):
    return np.loadtxt(path, ndmin=2)


def test_image_coordinates_and_write(tmp_path):
    """Image::set_coordinates (image.cpp:93-132) puts each cell on the plane
    normal to the ray: with the ray's unit vector n, ImX^2 + ImY^2 + (x.n)^2
    = |x|^2, and ImX is orthogonal to z; Image::write (image.cpp:48-86) uses
    Simulation/Image/{ImX,ImY,I_m,I_p}_<ray>."""
    m = _small()
    sim = to_simulation(m)
    r = 2
    img = core.Image(r, sim.parameters)
    img.set_coordinates(sim.geometry)
    x, y = np.array(img.ImX), np.array(img.ImY)
    n = np.asarray(m.rays[r])
    pos = np.asarray(m.position)
    np.testing.assert_allclose(x ** 2 + y ** 2 + (pos @ n) ** 2, np.sum(pos ** 2, axis=1), rtol=1e-12)
    d = np.hypot(n[0], n[1])
    np.testing.assert_allclose(x, (n[1] * pos[:, 0] - n[0] * pos[:, 1]) / d, rtol=1e-13)
    img.write(core.IoText(str(tmp_path) + "/"))
    base = str(tmp_path) + "/Simulation/Image/"
    np.testing.assert_array_equal(np.loadtxt(base + f"ImX_{r}.txt"), x)
    Im = _read_text_array(base + f"I_m_{r}.txt")
    assert Im.shape == (m.ncells, m.nfreqs) and not Im.any()


def test_image_without_image_discretisation_returns_minus_one(tmp_path):
    sim = to_simulation(_small())
    assert sim.compute_and_write_image(core.IoText(str(tmp_path) + "/"), 0) == -1


@pytest.mark.gpu
def test_compute_and_write_image_matches_oracle(tmp_path):
    from oracle.py_oracle import Oracle
    m = synthetic.envelope_model(n=8, seed=7)
    print("\n[host] envelope8 seed 7: compute_and_write_image vs oracle", flush=True)
    width = 3.0e-5
    orc = Oracle(m)
    orc.compute_spectral_discretisation_image(width)
    orc.compute_boundary_intensities()
    orc.compute_LTE_level_populations()
    sim = to_simulation(m)
    assert sim.compute_spectral_discretisation_image(width) == 0
    assert sim.compute_boundary_intensities() == 0
    assert sim.compute_LTE_level_populations() == 0
    io = core.IoText(str(tmp_path) + "/")
    assert sim.compute_level_populations(io, True, 5) == -1   # needs the line grid
    r = 3
    assert sim.compute_and_write_image(io, r) == 0
    base = str(tmp_path) + "/Simulation/Image/"
    Im_o, Ip_o = orc.compute_image(r)
    for name, ref in (("I_m", Im_o), ("I_p", Ip_o)):
        got = _read_text_array(base + f"{name}_{r}.txt")
        assert got.shape == ref.shape
        assert float(np.max(np.abs(got - ref) / np.abs(ref))) < 1e-10, name
    assert os.path.exists(base + f"ImY_{r}.txt")
    np.testing.assert_array_equal(np.array(sim.get_npoints_on_rays_rest())[r, :5],
                                  [len(orc.trace_ray(o, r, rest=True)[1]) +
                                   len(orc.trace_ray(o, int(orc.antipod()[r]), rest=True)[1]) + 1
                                   for o in range(5)])


@pytest.mark.gpu
def test_scattering_fills_radiation_u(tmp_path):
    """Parameters.use_scattering: Radiation::read sizes u [nrays/2][ncells*nfreqs]
    (radiation.cpp:117-126) and compute_radiation_field fills it."""
    from oracle.py_oracle import Oracle
    m = synthetic.envelope_model(n=6, seed=3)
    print("\n[host] envelope6 seed 3: scattering u", flush=True)
    sim = to_simulation(m, use_scattering=True)
    assert sim.compute_spectral_discretisation() == 0
    assert sim.compute_boundary_intensities() == 0
    assert sim.compute_LTE_level_populations() == 0
    sim.precision = core.FP64
    assert sim.compute_radiation_field() == 0
    u = np.array(sim.radiation.u)
    orc = Oracle(m).setup()
    orc.set_scattering(True)
    orc.compute_radiation_field()
    u_o = orc.u().reshape(u.shape)
    assert float(np.max(np.abs(u - u_o) / np.abs(u_o))) < 1e-10
    sim.write(core.IoText(str(tmp_path) + "/"))
    assert os.path.exists(str(tmp_path) + "/Radiation/u.txt")


@pytest.mark.gpu
def test_simulation_read_from_hdf5_matches_oracle(tmp_path):
    """A model written to HDF5 and read back into a Simulation (Model::read over
    IoPython("hdf5")) gives the oracle's radiation field."""
    if not hasattr(core, "IoPython"):
        pytest.skip("HDF5 C library not found at build time")
    from oracle.py_oracle import Oracle
    m = synthetic.envelope_model(n=6, seed=3)
    print("\n[host] envelope6 seed 3 through HDF5", flush=True)
    path = str(tmp_path) + "/model.hdf5"
    to_simulation(m).write(core.IoPython("hdf5", path))
    sim = core.Simulation()
    sim.read(core.IoPython("hdf5", path))
    assert sim.compute_spectral_discretisation() == 0
    assert sim.compute_boundary_intensities() == 0
    assert sim.compute_LTE_level_populations() == 0
    sim.precision = core.FP64
    assert sim.compute_radiation_field() == 0
    orc = Oracle(m).setup()
    orc.compute_radiation_field()
    J_o = orc.J().ravel()
    assert float(np.max(np.abs(np.array(sim.radiation.J) - J_o) / np.abs(J_o))) < 1e-10
