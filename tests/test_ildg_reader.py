"""The library's LIME/ILDG reader (nq_read_ildg_gauge_conf) on the reference's own
conf fixture, and its device plaquette (nq_gauge_plaquette).

The reader is pinned three ways: bit-equality with the independent Python reader
of the test fixtures, the file's own SciDAC checksum (scidac-checksum record,
checked by io/reader.hpp:31-47 after reading), and, on the GPU, the plaquette
the reference's test pins for this file, 0.6350055722288719 to 1e-14
(test/nissa/src/plaquette_computation.cpp:7-8)."""
import numpy as np
import pytest

import lime
import nissa_amd as nq
from helpers import tm_conf
from oracle import oracle as orc

L = (8, 4, 4, 4)


def test_reader_matches_python_reader_and_checksum(golden_dir):
    conf, read, comp = nq.read_ildg_gauge_conf(f"{golden_dir}/L4T8conf", L)
    ref = lime.read_ildg_conf(f"{golden_dir}/L4T8conf", L)
    assert np.array_equal(conf, ref)
    assert read != (0, 0) and read == comp


def test_reader_detects_corruption_and_bad_sizes(golden_dir, tmp_path):
    raw = bytearray(open(f"{golden_dir}/L4T8conf", "rb").read())
    i = raw.find(b"ildg-binary-data")
    assert i > 0
    raw[i + 128 + 1000] ^= 0x01  # one bit of the data of one site
    bad = tmp_path / "corrupt"
    bad.write_bytes(bytes(raw))
    _, read, comp = nq.read_ildg_gauge_conf(bad, L)
    assert read != comp
    with pytest.raises(nq.NqError) as e:
        nq.read_ildg_gauge_conf(f"{golden_dir}/L4T8conf", (8, 4, 4, 8))
    assert e.value.code == nq.NQ_EINVAL
    with pytest.raises(nq.NqError):
        nq.read_ildg_gauge_conf(tmp_path / "missing", L)


@pytest.mark.gpu
def test_device_plaquette_of_reference_conf(golden_dir):
    conf, _, _ = nq.read_ildg_gauge_conf(f"{golden_dir}/L4T8conf", L)
    with nq.Ctx(L) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        p = nq.gauge_plaquette(ctx, g)
        g.close()
    assert abs(p - 0.6350055722288719) < 1e-14


@pytest.mark.gpu
@pytest.mark.parametrize("Ls", [(8, 8, 8, 8), (4, 6, 4, 8), (10, 4, 2, 2)])
def test_device_plaquette_matches_oracle(Ls):
    conf = tm_conf(Ls, 7)  # antiperiodic phase on U_t: the plaquette does not see it
    ref = orc.Geom(Ls).plaquette(conf)
    with nq.Ctx(Ls) as ctx:
        g = nq.Gauge(ctx, lx=conf)
        p = nq.gauge_plaquette(ctx, g)
        g.close()
    assert abs(p - ref) < 1e-14


@pytest.mark.parametrize("fname", ["source_1.0003.00.inverted.00", "source_2.0003.00.inverted.04"])
def test_real_vector_reader_on_propagator_files(golden_dir, fname):
    """read_real_vector on the tmLQCD propagators of the reference's meson_2pts
    test: the same reals as the tests' reader and the file's own checksum."""
    v, read, comp = nq.read_lime_real_vector(f"{golden_dir}/{fname}", L)
    ref = lime.read_scidac_spincolor(f"{golden_dir}/{fname}", L)
    assert np.array_equal(v, ref.reshape(len(ref), 24))
    assert read != (0, 0) and read == comp


def test_unitarity_check(golden_dir):
    """nissa's post-read unitarity check: the reference conf is unitary to rounding,
    and a perturbed link gives the Frobenius norm of U U^+ - 1 that numpy computes."""
    conf, _, _ = nq.read_ildg_gauge_conf(f"{golden_dir}/L4T8conf", L)
    avg, mx, nb = nq.unitarity_check_lx(L, conf)
    assert mx < 1e-13 and avg <= mx and nb == 0
    U = conf[5, 2, ..., 0] + 1j * conf[5, 2, ..., 1]
    W = U @ (np.eye(3) + 1e-4 * np.array([[0, 1, 0], [0.5j, 0, 0], [0, 0, 2]]))
    conf[5, 2, ..., 0], conf[5, 2, ..., 1] = W.real, W.imag
    want = np.linalg.norm(W @ W.conj().T - np.eye(3))
    avg2, mx2, nb2 = nq.unitarity_check_lx(L, conf)
    assert abs(mx2 - want) < 1e-12 * want and nb2 == 1
    n = conf.shape[0] * 4  # links; the replaced link's own deviation was below 1e-13
    assert abs(avg2 - (avg * n + want) / n) < 1e-13
