"""
Autotuner logic without a device: candidate sets per coupling and
geometry, tuning validation, and the choice cache.
"""
import pytest

import myokit_amd as myokit
from myokit_amd import tune


def sim(ncells, **kw):
    return myokit.SimulationOpenCL(myokit.load_model('example'),
                                   myokit.pacing.blocktrain(1000, 2),
                                   ncells=ncells, **kw)


def test_candidates_by_coupling():
    c = tune.candidates(sim((256, 256)))
    drivers = {x['driver'] for x in c}
    assert drivers == {'block', 'grid', 'sheet'}
    sheet = [x for x in c if x['driver'] == 'sheet']
    assert {x['sheet_cpt'] for x in sheet} <= set(
        myokit_amd_choices()) and all(x['block_threads'] == 256
                                      for x in sheet)
    c = tune.candidates(sim(100))
    assert {x['driver'] for x in c} == {'block', 'grid'}
    c = tune.candidates(sim(100, diffusion=False))
    assert {x['driver'] for x in c} == {'cells', 'grid'}
    s = sim(10)
    s.set_connections([(i, i + 1, 1.0) for i in range(9)])
    c = tune.candidates(s)
    assert {x['driver'] for x in c} == {'block', 'grid', 'chain'}
    from myokit_amd import _kernelgen
    assert all(x['chain_cpt'] in _kernelgen.CHAIN_CPT_CHOICES
               for x in c if x['driver'] == 'chain')
    for x in tune.candidates(sim((64, 64), precision=64)):
        if x['driver'] == 'sheet':
            assert x['sheet_cpt'] in (2, 4)
    g64 = {x['grid_waves'] for x in tune.candidates(
        sim((64, 64), precision=64)) if x['driver'] == 'grid'}
    assert g64 == {0, 4, 5, 6}
    assert {x['grid_waves'] for x in tune.candidates(sim((64, 64)))
            if x['driver'] == 'grid'} == {0}


def myokit_amd_choices():
    from myokit_amd import _kernelgen
    return _kernelgen.SHEET_CPT_CHOICES


def test_set_tuning_validation():
    s = sim(50)
    assert s.tuning()['driver'] == 'auto'
    with pytest.raises(ValueError):
        s.set_tuning(block_threads=100)
    with pytest.raises(ValueError):
        s.set_tuning(sheet_cpt=3)
    with pytest.raises(ValueError):
        s.set_tuning(sheet_waves=0)
    with pytest.raises(ValueError):
        s.set_tuning(driver='fast')
    with pytest.raises(ValueError):
        s.set_tuning(adjoint_waves=9)
    with pytest.raises(ValueError):
        s.set_tuning(chain_cpt=3)
    with pytest.raises(ValueError):
        s.set_tuning(chain_waves=9)
    with pytest.raises(ValueError):
        s.set_tuning(adjoint_cpt=3)
    with pytest.raises(ValueError):
        s.set_tuning(grid_waves=9)
    s.set_tuning(driver='grid', block_threads=128, sheet_cpt=8,
                 sheet_waves=3, adjoint_waves=4, chain_cpt=1, chain_waves=6,
                 adjoint_cpt=4, grid_waves=5)
    t = s.tuning()
    assert t == {'driver': 'grid', 'block_threads': 128, 'sheet_cpt': 8,
                 'sheet_waves': 3, 'adjoint_waves': 4, 'chain_cpt': 1,
                 'chain_waves': 6, 'adjoint_cpt': 4, 'grid_waves': 5}


def test_cache_round_trip(tmp_path, monkeypatch):
    monkeypatch.setenv(tune.CACHE_ENV, str(tmp_path / 'tune.json'))
    s = sim((32, 32))
    assert tune.apply_cached(s) is None
    cfg = {'driver': 'grid', 'block_threads': 512}
    tune._store(tune._key(s), cfg)
    t = sim((32, 32))
    assert tune.apply_cached(t) == cfg
    assert t.tuning()['block_threads'] == 512
    # another geometry or precision has its own entry
    assert tune.apply_cached(sim((32, 16))) is None
    assert tune.apply_cached(sim((32, 32), precision=64)) is None


def test_tuned_kernel_source_has_tile_shape():
    from myokit_amd import _kernelgen
    from myokit_amd.simulation import _prepare_cell_model
    cm = _prepare_cell_model(myokit.load_model('example'), rl=True,
                             diffusion=True, fields=[], quantities=[])
    src = _kernelgen.generate_kernel(cm, 32, 1, tile_cpt=8, tile_waves=3)
    assert '#define MK_SCPT 8' in src and '#define MK_SWAVES 3' in src
    with pytest.raises(ValueError):
        _kernelgen.generate_kernel(cm, 32, 1, tile_cpt=16)


def test_grid_register_budget_in_kernel_source():
    # waves per SIMD -> workgroups per CU of the grid driver's launch
    # bounds (4 SIMDs per CU): 5 waves of 256-thread workgroups is 5 per
    # CU, of 1024-thread workgroups 1; 0 leaves the bound to the compiler
    from myokit_amd import _kernelgen
    from myokit_amd.simulation import _prepare_cell_model
    cm = _prepare_cell_model(myokit.load_model('example'), rl=True,
                             diffusion=True, fields=[], quantities=[])
    for threads, waves, blocks in ((256, 5, 5), (128, 4, 8), (1024, 5, 1),
                                   (256, 0, 0)):
        src = _kernelgen.generate_kernel(cm, 64, 1, threads,
                                         grid_budget=waves)
        assert '#define MK_GBLOCKS %d' % blocks in src
    with pytest.raises(ValueError):
        _kernelgen.generate_kernel(cm, 64, 1, grid_budget=9)
