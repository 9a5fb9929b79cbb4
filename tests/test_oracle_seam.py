"""The seam's read-out calls of the oracle agree with its full host copy of the wave packet, in the layouts the
reference's PropagationBackend contract fixes (src/propagation/propagation_backend.hpp:34-59): wave_packet_on_host
in (irot, ixy, izr) with izr = ir + nr * iz; z_plane(iz) in (irot, ixy, ir); block_densities() as sum |psi|^2 over
(irot, ixy) per izr; site_column(ixy) in (izr, irot). The first index of each layout runs fastest."""

import numpy as np
import pytest

from tests import cases
from tests.util import pseudo_random
from oracle import oracle


@pytest.fixture(scope="module")
def smoke():
    run = oracle.run(cases.write_case("seam_smoke", cases.SMOKE_DAILIGHT))
    s = pseudo_random(run.scattering_size, 5)
    p = pseudo_random(run.specular_size, 6)
    run.set_wave(s, p)
    run.propagate()
    yield run
    run.close()


def host_wave(run):
    i = run.info
    s, _ = run.get_wave()
    return s.reshape(i.nz * i.nr, i.nx * i.ny, i.n_rot)  # [izr][ixy][irot], irot fastest


def test_z_plane_is_the_host_copy_at_iz(smoke):
    i = smoke.info
    w = host_wave(smoke).reshape(i.nz, i.nr, i.nx * i.ny, i.n_rot)
    for iz in (0, i.iz_analysis, i.nz - 1):
        plane = smoke.z_plane(iz).reshape(i.nr, i.nx * i.ny, i.n_rot)
        np.testing.assert_array_equal(plane, w[iz])


def test_block_densities_are_sums_of_squares(smoke):
    w = host_wave(smoke)
    np.testing.assert_allclose(smoke.block_densities(), (np.abs(w) ** 2).sum(axis=(1, 2)), rtol=1e-13)


def test_site_column_is_the_host_copy_at_ixy(smoke):
    w = host_wave(smoke)
    for ixy in (0, 5, w.shape[1] - 1):
        column = smoke.site_column(ixy).reshape(w.shape[2], w.shape[0])  # [irot][izr], izr fastest
        np.testing.assert_array_equal(column.T, w[:, ixy, :])


def test_round_trip_of_set_and_get_wave():
    run = oracle.run(cases.write_case("seam_smoke", cases.SMOKE_DAILIGHT))
    s = pseudo_random(run.scattering_size, 9)
    p = pseudo_random(run.specular_size, 10)
    run.set_wave(s, p)
    s2, p2 = run.get_wave()
    np.testing.assert_array_equal(s, s2)
    np.testing.assert_array_equal(p, p2)
    run.close()
