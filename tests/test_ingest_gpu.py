"""The batched host -> HBM ingest (hdfe_stage_*): columns of several element
sizes arrive bit for bit, whatever the batch size, ring depth and thread
count (ragged last batches, one-row batches, empty input); Parquet files are
read batch by batch into the same arrays; a fit from a Parquet path or from
host arrays equals the fit from device tensors."""

import numpy as np
import pytest

from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs a GPU")]


def _cols(n, seed=0):
    rng = np.random.default_rng(seed)
    return {"a": rng.integers(-2**40, 2**40, n, dtype=np.int64),
            "b": rng.standard_normal(n),
            "c": rng.standard_normal(n).astype(np.float32),
            "d": rng.integers(0, 2**31 - 1, n).astype(np.int32),
            "e": rng.integers(0, 255, n).astype(np.uint8)}


@pytest.mark.parametrize("n,batch,depth,threads", [(12_345, 1_000, 2, 4), (7, 1, 1, 1),
                                                   (3_000_000, 1_000_000, 3, 16),
                                                   (1_000_000, 1_000_000, 2, 3)])
def test_columns_arrive_bit_for_bit(n, batch, depth, threads):
    import torch
    from hdfe_stream_amd.ingest import Stager
    print(f"stage n={n} batch={batch} depth={depth} threads={threads}", flush=True)
    cols = _cols(n)
    dev = [torch.full((n + 5,), 0, dtype=torch.from_numpy(v[:0]).dtype, device="cuda")
           for v in cols.values()]
    torch.cuda.synchronize()
    st = Stager([v.itemsize for v in cols.values()], batch_rows=batch, depth=depth,
                threads=threads)
    s = st.stage(list(cols.values()), dev, dev_row=5)
    st.close()
    assert s["batches"] == -(-n // batch)
    assert s["bytes"] == sum(v.nbytes for v in cols.values())
    for v, d in zip(cols.values(), dev):
        got = d.cpu().numpy()
        assert np.array_equal(got[5:], v) and not got[:5].any()


def test_to_device_casts_and_passes_device_tensors():
    import torch
    from hdfe_stream_amd.ingest import to_device
    cols = _cols(5000, seed=1)
    on_dev = torch.arange(5000, device="cuda")
    stats = {}
    out = to_device({"a": cols["a"], "d": cols["d"], "t": on_dev, "c": cols["c"]},
                    {"d": torch.int64, "t": torch.float64}, batch_rows=777, stats=stats)
    assert out["d"].dtype == torch.int64 and np.array_equal(out["d"].cpu().numpy(), cols["d"])
    assert out["t"].dtype == torch.float64 and out["t"].is_cuda
    assert out["c"].dtype == torch.float32
    assert np.array_equal(out["a"].cpu().numpy(), cols["a"])
    assert stats["bytes"] == 5000 * (8 + 8 + 4) and stats["GBps"] > 0
    assert to_device({"z": np.zeros(0)})["z"].numel() == 0


def test_parquet_is_read_batch_by_batch(tmp_path):
    import pyarrow as pa
    import pyarrow.parquet as pq
    import torch
    from hdfe_stream_amd.ingest import parquet_to_device
    cols = _cols(25_001, seed=2)
    path = tmp_path / "p.parquet"
    pq.write_table(pa.table(cols), path, row_group_size=4_000)
    stats = {}
    out = parquet_to_device(path, ["b", "a", "c"], {"a": torch.int64}, batch_rows=3_000,
                            stats=stats)
    for k in ("a", "b", "c"):
        assert np.array_equal(out[k].cpu().numpy(), cols[k]), k
    assert stats["bytes"] == 25_001 * 20


def test_fit_from_parquet_and_host_arrays_equals_device_fit(tmp_path, small_panel):
    import pyarrow as pa
    import pyarrow.parquet as pq
    import torch
    from hdfe_stream_amd.estimator import feols_stream
    fml = "y ~ x1 + x2 | worker_id + firm_id + year"
    keys = ("worker_id", "firm_id", "year", "x1", "x2", "y")
    d = {k: small_panel[k] for k in keys}
    path = tmp_path / "panel.parquet"
    pq.write_table(pa.table(d), path, row_group_size=1_000)
    ref = feols_stream(fml, {k: torch.from_numpy(v).cuda() for k, v in d.items()})
    for src in (d, str(path)):
        r = feols_stream(fml, src)
        assert np.array_equal(r.beta, ref.beta)
        assert np.array_equal(r.resid(), ref.resid())
        assert r.diagnostics["ingest"]["bytes"] == len(d["y"]) * 48
