"""Host side of the autotuner cache (gpu/autotune.py): size classes and the
JSON round trip."""
from pyscf_wb97mv_fast_amd.gpu.autotune import Tuner


def test_size_classes():
    assert Tuner.key(648, 218566, 2, 'mixed') == '1024/262144/2/mixed'
    assert Tuner.key(512, 262144, 2, 'mixed') == '512/262144/2/mixed'
    assert Tuner.key(513, 262145, 1, 'fp64') == '1024/524288/1/fp64'


def test_save_load(tmp_path):
    t = Tuner()
    t.table['1024/262144/2/mixed'] = {'ncopy': 4, 'fuse_mask': 5, 'best_ms': 1.0,
                                      'default_ms': 1.2, 'configs': 7}
    p = tmp_path / 't.json'
    t.save(str(p))
    assert Tuner().load(str(p)).table == t.table
