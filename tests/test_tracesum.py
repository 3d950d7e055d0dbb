"""Counter summaries of the profiling helper (CPU only, synthetic CSV files)."""
import csv

import pytest

from hprmat_amd import tracesum

HEAD = ['Correlation_Id', 'Dispatch_Id', 'Agent_Id', 'Queue_Id', 'Process_Id', 'Thread_Id',
        'Grid_Size', 'Kernel_Id', 'Kernel_Name', 'Workgroup_Size', 'LDS_Block_Size',
        'Scratch_Size', 'VGPR_Count', 'Accum_VGPR_Count', 'SGPR_Count', 'Counter_Name',
        'Counter_Value', 'Start_Timestamp', 'End_Timestamp']
KER = 'trailing_update_kernel<float, 64, 64, 16, 32, 32, true>'


def write(path, counter, rows):
    with open(path, 'w', newline='') as f:
        w = csv.writer(f)
        w.writerow(HEAD)
        for i, (name, wgs, value) in enumerate(rows):
            w.writerow([i, i + 1, 'Agent 1', 1, 1, 1, wgs * 256, 7, name + '(int, int)', 256,
                        0, 0, 64, 0, 32, counter, value, 1000 * i, 1000 * i + 500])


def test_factor_traffic_shapes_and_ratios(tmp_path):
    # N = 1024, nb = 256: look-ahead updates 768 x 512 and 512 x 256, then 256 x 256
    shapes = [(768, 512), (512, 256), (256, 256)]
    alg_r = [8 * (m * 256 + 256 * n + m * n) for m, n in shapes]
    alg_w = [8 * m * n for m, n in shapes]
    copy = 1 << 20
    # the copy reads half its bytes by FETCH_SIZE (scale 2) and all by WRITE_SIZE
    fetch = [('__amd_rocclr_copyBuffer', 256, copy / 1024 / 2)] + [
        (KER, (m // 64) * (n // 64), 1.5 * a / 1024 / 2) for (m, n), a in zip(shapes, alg_r)]
    wr = [('__amd_rocclr_copyBuffer', 256, copy / 1024)] + [
        (KER, (m // 64) * (n // 64), a / 1024) for (m, n), a in zip(shapes, alg_w)]
    write(tmp_path / 'f.csv', 'FETCH_SIZE', fetch)
    write(tmp_path / 'w.csv', 'WRITE_SIZE', wr)
    rec = tracesum.factor_traffic([str(tmp_path / 'f.csv'), str(tmp_path / 'w.csv')], copy, KER,
                                  1024, 256, 8)
    assert rec['launches'] == 3
    assert rec['fetch_scale'] == pytest.approx(2.0) and rec['write_scale'] == pytest.approx(1.0)
    assert rec['read_ratio'] == pytest.approx(1.5) and rec['write_ratio'] == pytest.approx(1.0)
    assert rec['alg_bytes'] == pytest.approx((sum(alg_r) + sum(alg_w)) / 3)
    assert rec['hbm_bytes'] == pytest.approx((1.5 * sum(alg_r) + sum(alg_w)) / 3)
    assert (rec['first_launch']['M'], rec['first_launch']['N']) == (768, 512)


def test_factor_traffic_rejects_unknown_grid(tmp_path):
    write(tmp_path / 'f.csv', 'FETCH_SIZE',
          [('__amd_rocclr_copyBuffer', 256, 512.0), (KER, 7, 100.0)])
    with pytest.raises(ValueError, match='fit no update'):
        tracesum.factor_traffic([str(tmp_path / 'f.csv')], 1 << 20, KER, 1024, 256, 8)
