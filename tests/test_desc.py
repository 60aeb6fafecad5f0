"""Descriptor validation of lmfhd_ctx_create (runs without a GPU: an invalid descriptor is
rejected before any device call): gmres_spatial_order 4 exists for visc_type 1 on periodic
boxes only, as the reference's stag_applyop_4o (stag_applyop.f90:1340: visc_type 1 only);
visc_schur_approx is 0 or 1 and 1 is built for |visc_type| 1 and 2; filtering_width is 0, 1, 2
or 4 (filter_stoch_m_flux, stochastic_m_fluxdiv.f90:669-670) on periodic boxes at least that wide."""
import pytest

from lowmachfhd_amd.grid import NO_SLIP_WALL, PERIODIC, Grid


@pytest.mark.parametrize("kw", [
    dict(visc_type=-1, gmres_spatial_order=4),
    dict(visc_type=1, gmres_spatial_order=3),
    dict(visc_type=1, visc_schur_approx=2),
    dict(visc_type=3, visc_schur_approx=1),
    dict(visc_type=1, precon_type=6),
    dict(visc_type=1, filtering_width=3),
    dict(visc_type=1, filtering_width=-1),
    dict(visc_type=0),
])
def test_invalid_desc_rejected(kw):
    from lowmachfhd_amd import solver as S

    with pytest.raises(S.LmfhdError, match="E_ARG"):
        S.StokesSolver(Grid((16, 16, 16), 1 / 16), **kw)


def test_order4_rejects_walls():
    from lowmachfhd_amd import solver as S

    gw = Grid((16, 16, 16), 1 / 16, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL))
    with pytest.raises(S.LmfhdError, match="E_ARG"):
        S.StokesSolver(gw, visc_type=1, gmres_spatial_order=4)


@pytest.mark.parametrize("g", [
    Grid((16, 16, 16), 1 / 16, bc_lo=(PERIODIC, PERIODIC, NO_SLIP_WALL), bc_hi=(PERIODIC, PERIODIC, NO_SLIP_WALL)),
    Grid((16, 16, 2), 1 / 16),
])
def test_filtering_needs_periodic_boxes_at_least_width(g):
    from lowmachfhd_amd import solver as S

    with pytest.raises(S.LmfhdError, match="E_ARG"):
        S.StokesSolver(g, visc_type=1, filtering_width=4)
