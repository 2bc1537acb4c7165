"""The C++ host layer: .ini grammar and derived constants (CPU), and the -p
workflow on the GPU against the oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cosmicatlas_amd", "bin", "cosmicatlas")


def _ini(tmp_path, **kv):
    base = dict(Nft="128", Lbox="500.0", mass_assignment_scheme="CIC", MAS_correction="true",
                type_of_binning="linear", DeltaKmin="1", ndel_data="1", SN_correction="true",
                Statistics="Pk_fkp", Input_type="catalog", Output_directory=str(tmp_path) + "/",
                file_power="test", Name_survey="BOX", Input_dir_cat=str(tmp_path) + "/",
                Catalogue_file="cat.txt", i_coord1_g="0", i_coord2_g="1", i_coord3_g="2")
    base.update(kv)
    lines = ["# comment line", "Unknown_key = 3   # ignored like the reference"]
    lines += [f"{k} = {v}   # trailing text after the value is ignored" for k, v in base.items()]
    p = tmp_path / "run.ini"
    p.write_text("\n".join(lines) + "\n")
    return str(p)


def _show(ini):
    r = subprocess.run([CLI, "-i", ini], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return dict(line.split(" = ", 1) for line in r.stdout.strip().splitlines())


def test_derived_parameters(tmp_path):
    d = _show(_ini(tmp_path))
    kf = 2 * math.pi / 500.0
    assert int(d["Nft"]) == 128 and int(d["NGRID"]) == 128 ** 3 and int(d["NGRID_h"]) == 128 * 128 * 65
    assert float(d["delta_x"]) == 500.0 / 128
    assert math.isclose(float(d["deltak_0"]), kf, rel_tol=1e-15)
    assert math.isclose(float(d["kmin"]), kf, rel_tol=1e-15)
    assert math.isclose(float(d["DeltaK_data"]), kf, rel_tol=1e-15)
    assert int(d["Nnp_data"]) == 64 and int(d["mass_assignment"]) == 1
    d2 = _show(_ini(tmp_path, Nft="512", ndel_data="2", mass_assignment_scheme="TSC"))
    assert int(d2["Nnp_data"]) == 128 and int(d2["mass_assignment"]) == 2


@pytest.mark.parametrize("line,msg", [("Nft 128", "'=' not properly written"),
                                      ("mass_assignment_scheme = QQQ", "Unknown mass assigment scheme"),
                                      ("type_of_binning = cubic", "Unknown type of binning")])
def test_ini_errors_exit_1(tmp_path, line, msg):
    p = tmp_path / "bad.ini"
    p.write_text("Nft = 64\n" + line + "\n")
    r = subprocess.run([CLI, "-i", str(p)], capture_output=True, text=True)
    assert r.returncode == 1 and msg in r.stderr


def test_missing_file_and_usage():
    r = subprocess.run([CLI, "-i", "/nonexistent.ini"], capture_output=True, text=True)
    assert r.returncode == 1 and "Error at opening the parameter file" in r.stderr
    assert subprocess.run([CLI], capture_output=True).returncode == 1


def test_empty_value_keeps_default(tmp_path):
    p = tmp_path / "e.ini"
    p.write_text("Nft = 64\nLbox =\n")
    d = _show(str(p))
    assert float(d["Lbox"]) == 100.0


@pytest.mark.gpu
def test_cli_catalog_power_spectrum(tmp_path, orc):
    N, L, n = 64, 500.0, 100000
    pos = np.random.default_rng(21).uniform(0, L, (n, 3)).astype(np.float32)
    np.savetxt(tmp_path / "cat.txt", pos, fmt="%.7f")
    ini = _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", compute_tweb="true", smscale="20",
               lambdath="0.0")
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.loadtxt(tmp_path / f"Pk_fkp_BOX_Nft{N}_TSC_test.txt")
    pos32 = np.loadtxt(tmp_path / "cat.txt").astype(np.float32)
    mesh = orc.mas(pos32, N, L, orc.TSC)
    V = L ** 3
    kf = 2 * np.pi / L
    dk = orc.fft_r2c(mesh - n / N ** 3)
    o = orc.pk_opts(N, L, N // 2, kf, kf, orc.TSC, V / n, n * n / V)
    p0, _, _, nm = orc.pk_shells(dk, o)
    k = kf + kf * (np.arange(N // 2) + 0.5)
    first = 1 if p0[0] < 1 else 0
    # the k P N text has the stream's default 6 significant digits (FileOutput.cpp:490-511)
    np.testing.assert_allclose(out[:, 0], k[first:], rtol=1e-5)
    assert np.array_equal(out[:, 2].astype(int), nm[first:])
    assert np.max(np.abs(out[:, 1] - p0[first:]) / (np.abs(p0[first:]) + V / n)) < 1e-4
    labels = np.fromfile(tmp_path / f"CWC_BOX_Nft{N}_TSC_test.dat", np.float32)
    assert labels.size == N ** 3 and set(np.unique(labels)) <= {1, 2, 3, 4}


@pytest.mark.gpu
def test_cli_log_binning(tmp_path, orc):
    """type_of_binning = log through the -p workflow: N_log_bins shells with
    the reference's rule floor(log10(|k|/kmin)/Deltal) + 1 (FftwFunctions.cpp:
    1103-1108, Params.cpp:4104-4106), k written at the log centre
    kmin 10^((s - 1/2) Deltal)."""
    N, L, n, nlog = 64, 500.0, 100000, 8
    pos = np.random.default_rng(23).uniform(0, L, (n, 3)).astype(np.float32)
    np.savetxt(tmp_path / "cat.txt", pos, fmt="%.7f")
    ini = _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", type_of_binning="log",
               N_log_bins=str(nlog), DeltaKmin="1")
    d = _show(ini)
    kf = 2 * np.pi / L
    dlog = math.log10((N / 2) * kf / kf) / nlog
    assert int(d["Nnp_data"]) == nlog
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.loadtxt(tmp_path / f"Pk_fkp_BOX_Nft{N}_TSC_test.txt")
    pos32 = np.loadtxt(tmp_path / "cat.txt").astype(np.float32)
    mesh = orc.mas(pos32, N, L, orc.TSC)
    V = L ** 3
    dk = orc.fft_r2c(mesh - n / N ** 3)
    o = orc.pk_opts(N, L, nlog, kf, 0.0, orc.TSC, V / n, n * n / V, log_binning=True, dlog=dlog)
    p0, _, _, nm = orc.pk_shells(dk, o)
    k = kf * 10.0 ** ((np.arange(nlog) - 0.5) * dlog)
    first = 1 if p0[0] < 1 else 0
    np.testing.assert_allclose(out[:, 0], k[first:], rtol=1e-5)
    assert np.array_equal(out[:, 2].astype(int), nm[first:])
    assert np.max(np.abs(out[:, 1] - p0[first:]) / (np.abs(p0[first:]) + V / n)) < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("nbar_cols,fkp", [(False, True), (True, True), (False, False)])
def test_cli_random_catalogue(tmp_path, orc, nbar_cols, fkp):
    """use_random_catalog through the -p workflow (PowerSpectrumF.cpp:741-1363
    random branch; FftwFunctions.cpp:177-200, 275-281): per-object sums with
    FKP weights 1/(1 + Pest nbar) (Catalog.cpp:14142-14171), alpha = W_g/W_r,
    delta = rho_g - alpha rho_r, shot noise (alpha + 1) S_r/normal_r, and the
    window file of the randoms' spectrum over normal_power/alpha^2.  Against the
    oracle's deposits, FFT, shells and window mode set."""
    N, L, scheme = 32, 200.0, "CIC"
    rng = np.random.default_rng(41)
    ng, nr = 8000, 30000
    gal = rng.uniform(0, L, (ng, 3))
    gal[:, 0] = L * rng.beta(1.5, 1.5, ng)  # a density gradient along x
    ran = rng.uniform(0, L, (nr, 3))
    wg = rng.uniform(0.5, 1.5, ng)
    wr = rng.uniform(0.5, 1.5, nr)
    nbg = 1e-3 * (1.0 + gal[:, 0] / L)
    nbr = 1e-3 * (1.0 + ran[:, 0] / L)
    np.savetxt(tmp_path / "cat.txt", np.column_stack([gal, wg, nbg]), fmt="%.8g")
    np.savetxt(tmp_path / "ran.txt", np.column_stack([ran, wr, nbr]), fmt="%.8g")
    Pest = 2000
    kv = dict(Nft=str(N), Lbox=str(L), mass_assignment_scheme=scheme, DeltaKmin="0", use_random_catalog="true",
              Random_file="ran.txt", i_coord1_r="0", i_coord2_r="1", i_coord3_r="2", i_weight1_g="3",
              i_weight1_r="3", FKP_weight="true" if fkp else "false", Pest=str(Pest), file_window="win")
    if nbar_cols:
        kv.update(i_mean_density_g="4", i_mean_density_r="4")
    ini = _ini(tmp_path, **kv)
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.loadtxt(tmp_path / f"Pk_fkp_BOX_Nft{N}_{scheme}_test.txt")
    win = np.loadtxt(tmp_path / f"Pk_fkp_BOX_Nft{N}_{scheme}_win.txt")
    # the catalogue values as the reader sees them (float32 columns)
    G = np.loadtxt(tmp_path / "cat.txt").astype(np.float32).astype(np.float64)
    R = np.loadtxt(tmp_path / "ran.txt").astype(np.float32).astype(np.float64)
    V = L ** 3
    nbar_c = G[:, 3].sum() / V
    ng_, nr_ = (G[:, 4], R[:, 4]) if nbar_cols else (np.full(ng, nbar_c), np.full(nr, nbar_c))
    fg = G[:, 3] / (1.0 + Pest * ng_) if fkp else G[:, 3]
    fr = R[:, 3] / (1.0 + Pest * nr_) if fkp else R[:, 3]
    alpha = fg.sum() / fr.sum()
    normal_r = (nr_ * fr ** 2).sum()
    SN = (alpha + 1) * (fr ** 2).sum() / normal_r
    mr = orc.mas(R[:, :3].astype(np.float32), N, L, orc.CIC, weights=fr.astype(np.float32))
    mesh = orc.mas(G[:, :3].astype(np.float32), N, L, orc.CIC, weights=fg.astype(np.float32)) - alpha * mr
    kf = 2 * np.pi / L
    p0, _, _, nm = orc.pk_shells(orc.fft_r2c(mesh), orc.pk_opts(N, L, N // 2, 0.0, kf, orc.CIC, SN, alpha * normal_r))
    first = 1 if p0[0] < 1 else 0
    assert np.array_equal(out[:, 2].astype(int), nm[first:])
    assert np.max(np.abs(out[:, 1] - p0[first:]) / (np.abs(p0[first:]) + SN)) < 1e-4
    pw, _ = orc.window_shells(orc.fft_r2c(mr), orc.pk_opts(N, L, N // 2, 0.0, kf, orc.CIC, 0.0,
                                                             normal_r / alpha))
    firstw = 1 if pw[0] < 1 else 0
    assert np.array_equal(win[:, 2].astype(int), nm[firstw:])  # the data shells' counts, as the reference writes
    np.testing.assert_allclose(win[:, 1], pw[firstw:], rtol=2e-4)


@pytest.mark.gpu
def test_cli_density_grid(tmp_path, orc):
    N, L = 32, 300.0
    rng = np.random.default_rng(5)
    grid = rng.poisson(3.0, (N, N, N)).astype(np.float32)
    grid.tofile(tmp_path / "grid.dat")
    ini = _ini(tmp_path, Nft=str(N), Lbox=str(L), Input_type="density_grid", mass_assignment_scheme="NGP",
               delta_grid_file=str(tmp_path / "grid.dat"), DeltaKmin="0")
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.loadtxt(tmp_path / f"Pk_fkp_BOX_Nft{N}_NGP_test.txt")
    ng = grid.astype(np.float64).sum()
    delta = grid / (ng / N ** 3) - 1.0
    dk = orc.fft_r2c(delta)
    kf = 2 * np.pi / L
    o = orc.pk_opts(N, L, N // 2, 0.0, kf, orc.NGP, L ** 3 / ng, (N ** 3 / L ** 1.5) ** 2)
    p0, _, _, nm = orc.pk_shells(dk, o)
    first = 1 if p0[0] < 1 else 0
    assert np.array_equal(out[:, 2].astype(int), nm[first:])
    assert np.max(np.abs(out[:, 1] - p0[first:]) / (np.abs(p0[first:]) + L ** 3 / ng)) < 1e-4


def _run_pk(tmp_path, ini, N):
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return np.loadtxt(tmp_path / f"Pk_fkp_BOX_Nft{N}_TSC_test.txt")


def _oracle_pk(orc, pos32, N, L):
    n = len(pos32)
    mesh = orc.mas(pos32, N, L, orc.TSC)
    V = L ** 3
    kf = 2 * np.pi / L
    dk = orc.fft_r2c(mesh - n / N ** 3)
    p0, _, _, nm = orc.pk_shells(dk, orc.pk_opts(N, L, N // 2, kf, kf, orc.TSC, V / n, n * n / V))
    first = 1 if p0[0] < 1 else 0
    return p0[first:], nm[first:], V / n


@pytest.mark.gpu
def test_cli_binary_catalogue(tmp_path, orc):
    """Catalog::read_catalog_bin: three headerless float32 coordinate files."""
    N, L, n = 64, 500.0, 60000
    pos = np.random.default_rng(22).uniform(0, L, (n + 500, 3)).astype(np.float32)
    for a, ax in enumerate("xyz"):
        pos[:, a].tofile(tmp_path / f"{ax}.dat")
    ini = _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", N_lines_binary=str(n),
               file_bin_x_coord=str(tmp_path / "x.dat"), file_bin_y_coord=str(tmp_path / "y.dat"),
               file_bin_z_coord=str(tmp_path / "z.dat"))
    out = _run_pk(tmp_path, ini, N)
    p0, nm, sn = _oracle_pk(orc, pos[:n], N, L)
    assert np.array_equal(out[:, 2].astype(int), nm)
    assert np.max(np.abs(out[:, 1] - p0) / (np.abs(p0) + sn)) < 1e-4


@pytest.mark.gpu
def test_cli_equatorial_catalogue(tmp_path, orc):
    """Equatorial (RA, Dec, distance) columns in degrees and in radians give the
    Cartesian measurement (equatorial_to_cartesian, CoordinateSystem.cpp:12-18)."""
    N, L, n = 64, 500.0, 60000
    rng = np.random.default_rng(23)
    xyz = rng.uniform(20.0, 480.0, (n, 3))
    r = np.sqrt((xyz ** 2).sum(1))
    dec = np.degrees(np.arcsin(xyz[:, 2] / r))
    ra = np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))
    np.savetxt(tmp_path / "cat.txt", np.stack([ra, dec, r], 1), fmt="%.12f")
    ini = _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", sys_of_coord_g="2", angles_units_g="D")
    deg = _run_pk(tmp_path, ini, N)
    np.savetxt(tmp_path / "cat.txt", np.stack([np.radians(ra), np.radians(dec), r], 1), fmt="%.14f")
    ini = _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", sys_of_coord_g="2", angles_units_g="R")
    rad = _run_pk(tmp_path, ini, N)
    p0, nm, sn = _oracle_pk(orc, xyz.astype(np.float32), N, L)
    for out in (deg, rad):
        assert np.array_equal(out[:, 2].astype(int), nm)
        assert np.max(np.abs(out[:, 1] - p0) / (np.abs(p0) + sn)) < 1e-4


C_KMS = 299792.458


def _rz_table(zmax, nz, om=0.3089, orad=1.3157e-5 * (1 + 0.2271 * 3.046), ok=0.0, ov=None, w=-1.0):
    """r(z) nodes by scipy's adaptive quadrature of c/H (independent of the host's Simpson rule)."""
    from scipy.integrate import quad
    ov = 1 - om - ok - orad if ov is None else ov
    hinv = lambda z: 1.0 / (100.0 * math.sqrt(om * (1 + z) ** 3 + orad * (1 + z) ** 4 + ok * (1 + z) ** 2
                                               + ov * (1 + z) ** (3 * (1 + w))))
    zc = np.arange(nz) * zmax / (nz - 1)  # z_i = i zmax/(nz-1), Source/Catalog.cpp:31
    rc = np.array([C_KMS * quad(hinv, 0.0, z, epsabs=0, epsrel=1e-13, limit=200)[0] for z in zc])
    return zc, rc


def _table(tmp_path, **kv):
    d = _show(_ini(tmp_path, sys_of_coord_g="1", **kv))
    return np.array(d["zcosmo"].split(), float), np.array(d["rcosmo"].split(), float)


def test_redshift_table_closed_forms(tmp_path):
    """Comoving distance against the two closed forms: Einstein-de Sitter,
    r = 2c/H0 (1 - (1+z)^-1/2), and de Sitter, r = c z/H0."""
    z, r = _table(tmp_path, om_matter="1", om_radiation="0", om_vac="0", om_k="0", Nbins_redshift="9",
                  redshift_max_sample="4")
    assert np.array_equal(z, np.arange(9) * 4.0 / 8)
    np.testing.assert_allclose(r, 2 * C_KMS / 100 * (1 - 1 / np.sqrt(1 + z)), rtol=1e-12, atol=0)
    z, r = _table(tmp_path, om_matter="0", om_radiation="0", om_vac="1", om_k="0", Nbins_redshift="5",
                  redshift_max_sample="2")
    np.testing.assert_allclose(r, C_KMS * z / 100, rtol=1e-13, atol=0)


@pytest.mark.parametrize("kv", [{}, dict(om_matter="0.25", om_vac="0.68", om_k="0.07", wde_eos="-0.8",
                                         om_radiation="9e-5", Nbins_redshift="13", redshift_max_sample="3")])
def test_redshift_table_quadrature(tmp_path, kv):
    """The r(z) table (Catalog::get_cosmo, Source/Catalog.cpp:16-37) against
    scipy's quadrature, for the default Planck background and a curved w != -1 one;
    the default r(1) is the textbook 2300.5 Mpc/h for Omega_m = 0.3089."""
    z, r = _table(tmp_path, **kv)
    args = {} if not kv else dict(om=0.25, orad=9e-5, ok=0.07, ov=0.68, w=-0.8)
    zc, rc = _rz_table(float(kv.get("redshift_max_sample", 1)), int(kv.get("Nbins_redshift", 10)), **args)
    assert np.array_equal(z, zc)
    np.testing.assert_allclose(r, rc, rtol=1e-11, atol=0)
    if not kv:
        assert abs(r[-1] - 2300.5) < 0.1


def test_redshift_above_table_is_an_error(tmp_path):
    """Redshifts beyond the last table node stop the run (GSL's interpolation
    refuses them); the catalogue is converted before any device work."""
    (tmp_path / "cat.txt").write_text("10 20 0.5\n30 -10 1.5\n")
    ini = _ini(tmp_path, Nft="32", sys_of_coord_g="1")
    rr = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=120)
    assert rr.returncode == 1 and "above the r(z) table" in rr.stderr + rr.stdout


@pytest.mark.gpu
def test_cli_redshift_catalogue(tmp_path, orc):
    """(RA, Dec, redshift) columns (Catalog::ang_to_cart_coordinates, I_EQZ
    branch, Source/Catalog.cpp:13786-13820): redshifts chosen so that linear
    interpolation in the r(z) table returns the distances of known Cartesian
    positions give those positions' measurement; redshifts below the first node
    are raised to it (distance 0)."""
    N, L, n = 64, 500.0, 60000
    rng = np.random.default_rng(24)
    xyz = rng.uniform(20.0, 480.0, (n, 3))
    r = np.sqrt((xyz ** 2).sum(1))
    zc, rc = _rz_table(1.0, 10)
    i = np.searchsorted(rc, r, side="right") - 1
    zz = zc[i] + (r - rc[i]) / (rc[i + 1] - rc[i]) * (zc[i + 1] - zc[i])
    dec = np.degrees(np.arcsin(xyz[:, 2] / r))
    ra = np.degrees(np.arctan2(xyz[:, 1], xyz[:, 0]))
    rows = np.stack([ra, dec, zz], 1)
    rows = np.vstack([rows, [[10.0, 20.0, -0.01]]])  # clamped to z = 0: the origin
    np.savetxt(tmp_path / "cat.txt", rows, fmt="%.12f")
    out = _run_pk(tmp_path, _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", sys_of_coord_g="1"), N)
    ref = np.vstack([xyz, [[0.0, 0.0, 0.0]]]).astype(np.float32)
    p0, nm, sn = _oracle_pk(orc, ref, N, L)
    assert np.array_equal(out[:, 2].astype(int), nm)
    assert np.max(np.abs(out[:, 1] - p0) / (np.abs(p0) + sn)) < 1e-4


def _hubble(z, om=0.3089, orad=1.3157e-5 * (1 + 0.2271 * 3.046)):
    return 100.0 * math.sqrt(om * (1 + z) ** 3 + orad * (1 + z) ** 4 + (1 - om - orad))


@pytest.mark.parametrize("units,z,factor", [
    ("kmps", 0.0, lambda z: 1.0 / _hubble(z)),              # 1/H0 = 0.01 Mpc/h per km/s
    ("kmps", 0.57, lambda z: (1 + z) / _hubble(z)),
    ("alpt", 1.2, lambda z: 1.0 / _hubble(z)),
    ("Mpcph", 0.57, lambda z: 1.0),
    ("other", 0.57, lambda z: 1.0)])
def test_rsd_conversion_factor(tmp_path, units, z, factor):
    """Velocity to line-of-sight displacement (Catalog.cpp:14088-14099) against
    (1+z)/H(z) written out for the default flat background."""
    d = _show(_ini(tmp_path, i_v3_g="5", vel_units_g=units, Redshift=str(z)))
    assert d["redshift_space_coords_g"] == "false"
    assert float(d["rsd_conversion_factor"]) == pytest.approx(factor(z), rel=1e-7)
    if units == "kmps" and z == 0.0:
        assert float(d["rsd_conversion_factor"]) == pytest.approx(0.01, rel=1e-7)


def test_velocity_column_rules(tmp_path):
    """A velocity column is read only at a positive index (Catalog.cpp:2306,
    `if(i_vz>0)`), and redshift_space_coords_g is shown with it."""
    assert "rsd_conversion_factor" not in _show(_ini(tmp_path, i_v3_g="0"))
    assert "rsd_conversion_factor" not in _show(_ini(tmp_path))
    d = _show(_ini(tmp_path, i_v3_g="4", redshift_space_coords_g="true"))
    assert d["redshift_space_coords_g"] == "true"


@pytest.mark.gpu
@pytest.mark.parametrize("coords_in_rsd", [False, True])
def test_cli_redshift_space_catalogue(tmp_path, orc, coords_in_rsd):
    """x y z vx vy vz columns: in real-space coordinates every object moves by
    vz (1+z)/H(z) along the third axis (Catalog::get_interpolated_density_field,
    Source/Catalog.cpp:14071-14129), wrapping through the box; with
    redshift_space_coords_g the velocities are ignored. Column 0 is never a
    velocity (the reference reads one when its index is positive)."""
    N, L, n, z = 64, 500.0, 60000, 0.57
    rng = np.random.default_rng(25)
    pos = rng.uniform(0, L, (n, 3)).astype(np.float32)
    vel = rng.normal(0, 600.0, (n, 3)).astype(np.float32)  # km/s: up to ~25 Mpc/h, many wrap
    np.savetxt(tmp_path / "cat.txt", np.hstack([pos, vel]), fmt="%.7f")
    rows = np.loadtxt(tmp_path / "cat.txt").astype(np.float32)
    pos, vel = rows[:, :3].copy(), rows[:, 3:]
    ini = _ini(tmp_path, Nft=str(N), mass_assignment_scheme="TSC", i_v1_g="3", i_v2_g="4", i_v3_g="5",
               Redshift=str(z), redshift_space_coords_g="true" if coords_in_rsd else "false")
    out = _run_pk(tmp_path, ini, N)
    ref = pos.copy()
    if not coords_in_rsd:
        f = np.float32(float(_show(ini)["rsd_conversion_factor"]))
        assert f == pytest.approx((1 + z) / _hubble(z), rel=1e-7)
        ref[:, 2] = pos[:, 2] + f * vel[:, 2]  # float32 like the reference's real_prec
        assert (ref[:, 2] < 0).sum() > 100 and (ref[:, 2] >= L).sum() > 100
    p0, nm, sn = _oracle_pk(orc, ref, N, L)
    assert np.array_equal(out[:, 2].astype(int), nm)
    assert np.max(np.abs(out[:, 1] - p0) / (np.abs(p0) + sn)) < 1e-4
    if not coords_in_rsd:
        # the displacement matters at this tolerance: the real-space spectrum differs
        p_real, _, _ = _oracle_pk(orc, pos, N, L)
        assert np.max(np.abs(out[:, 1] - p_real) / (np.abs(p_real) + sn)) > 1e-2


def _bk_catalogue(tmp_path, N, L, n, seed):
    pos = np.random.default_rng(seed).uniform(0, L, (n, 3)).astype(np.float32)
    np.savetxt(tmp_path / "cat.txt", pos, fmt="%.7f")
    return np.loadtxt(tmp_path / "cat.txt").astype(np.float32)


@pytest.mark.gpu
def test_cli_direct_bispectrum(tmp_path, orc):
    """Statistics = Bk_fkp (FftwFunctions::get_bispectrum_fkp): rows a b c k_a k_b
    k_c B SN_sum ntri for every bin triple with triangles, against the oracle's
    direct sums on the oracle's own mesh."""
    N, L, n = 16, 300.0, 20000
    pos = _bk_catalogue(tmp_path, N, L, n, 31)
    ini = _ini(tmp_path, Nft=str(N), Lbox=str(L), mass_assignment_scheme="CIC", Statistics="Bk_fkp",
               file_bispectrum="bk")
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = np.loadtxt(tmp_path / f"Bk_fkp_BOX_Nft{N}_CIC_bk.txt")
    mesh = orc.mas(pos, N, L, orc.CIC)
    dk = orc.fft_r2c(mesh - n / N ** 3)
    nb = N // 2
    b, sn, cnt, _, _ = orc.bk_direct(dk, 1, 1, nb, 1, True)
    a_, b_, c_ = rows[:, 0].astype(int), rows[:, 1].astype(int), rows[:, 2].astype(int)
    assert np.array_equal(np.sort(np.ravel_multi_index((a_, b_, c_), (nb,) * 3)),
                          np.sort(np.ravel_multi_index(np.nonzero(cnt), (nb,) * 3)))
    assert np.array_equal(rows[:, 8].astype(np.int64), cnt[a_, b_, c_])
    kf = 2 * np.pi / L
    np.testing.assert_allclose(rows[:, 3], kf + kf * (a_ + 0.5), rtol=1e-6)
    V = L ** 3
    m = cnt[a_, b_, c_]
    snb = (sn[a_, b_, c_] / m / (n * n / V) - 3 * V / n) * V / n + (V / n) ** 2
    B = b[a_, b_, c_] / (m * n ** 3 / V ** 2) - snb
    scale = (sn[a_, b_, c_] / m / 3) ** 1.5 / (n ** 3 / V ** 2) / np.sqrt(m) + np.abs(snb)
    assert np.max(np.abs(rows[:, 6] - B) / scale) < 1e-3
    np.testing.assert_allclose(rows[:, 7], sn[a_, b_, c_], rtol=1e-5)


@pytest.mark.gpu
def test_cli_fast_bispectrum(tmp_path, orc):
    """Statistics = Bk_fkp_fast: rows i j k B SN ntri (i >= j >= k) of the shells
    DeltaK_data wide up to kmax_bk, from kmin_bk; without shot noise B equals the
    direct estimator's and ntri x (distinct orderings) its pair count."""
    import itertools
    N, L, n = 32, 300.0, 50000
    kf = 2 * np.pi / L
    pos = _bk_catalogue(tmp_path, N, L, n, 32)
    ini = _ini(tmp_path, Nft=str(N), Lbox=str(L), mass_assignment_scheme="TSC", Statistics="Bk_fkp_fast",
               SN_correction="false", kmax_bk=str(9.5 * kf), kmin_bk=str(3.2 * kf), file_bispectrum="bk")
    r = subprocess.run([CLI, "-p", ini], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    rows = np.atleast_2d(np.loadtxt(tmp_path / f"Bk_fkp_fast_BOX_Nft{N}_TSC_bk.txt"))
    ns = 9
    # the outer shell starts at floor((3.2 - 1)/1) = 2; the inner loops of
    # loop_shells_bispectrum (FftwFunctions.cpp:3530-3541) may go below it
    assert rows[:, 0].min() == 2 and rows[:, :3].min() >= 0 and rows[:, :3].max() == ns - 1
    assert np.all((rows[:, 0] >= rows[:, 1]) & (rows[:, 1] >= rows[:, 2]))
    mesh = orc.mas(pos, N, L, orc.TSC)
    b, _, cnt, _, _ = orc.bk_direct(orc.fft_r2c(mesh - n / N ** 3), 1, 1, ns, 2, True)
    V = L ** 3
    checked = 0
    for i, j, k, B, sn, nt in rows:
        a, bb, c = sorted((int(i), int(j), int(k)))
        perms = len(set(itertools.permutations((a, bb, c))))
        assert cnt[a, bb, c] == perms * int(nt)
        assert sn == 0
        if nt:
            ref = b[a, bb, c] / (float(cnt[a, bb, c]) * float(n) ** 3 / V ** 2)
            # float32 shell fields: an error floor of a few 1e-7 of the largest bin
            assert abs(B - ref) <= 2e-4 * abs(ref) + 2e-6 * np.abs(rows[:, 3]).max(), (i, j, k, B, ref)
            checked += 1
    assert checked > 20
