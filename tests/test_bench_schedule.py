"""The benchmark's in-place step scheduler and argument handling (no GPU).

AA streaming alternates an ODD kernel (first) and an EVEN kernel on one array; the benchmark splits a run into
warm-up, timed and untimed-tail launches.  One global counter must drive the parity across all three, and the
run must end after an EVEN kernel so the final state is read in the layout an EVEN kernel leaves."""
import itertools

import pytest

import bench


@pytest.mark.parametrize("warmup,steps,total", list(itertools.product([0, 1, 4, 5], [1, 2, 19, 20], [0, 7, 2000])))
def test_schedule_is_one_contiguous_run_ending_even(warmup, steps, total):
    warm, timed, tail = bench.aa_schedule(warmup, steps, total)
    seq = list(warm) + list(timed) + list(tail)
    assert seq == list(range(len(seq)))                # one counter: kernel t % 2 for every launch
    assert len(warm) == warmup and len(timed) == steps
    assert len(seq) % 2 == 0                           # the last launch is an EVEN kernel
    assert len(seq) >= max(total, warmup + steps) and len(seq) - max(total, warmup + steps) <= 1


def test_odd_warmup_keeps_the_parity_of_the_timed_launches():
    # the driver's invocation: --steps 20 --warmup 5 -> the first timed launch is global step 5 (EVEN kernel)
    warm, timed, tail = bench.aa_schedule(5, 20, 2000)
    assert timed[0] == 5 and timed[0] % 2 == 1
    assert tail.stop == 2000


def test_schedule_rejects_bad_counts():
    with pytest.raises(ValueError):
        bench.aa_schedule(0, 0, 10)
    with pytest.raises(ValueError):
        bench.aa_schedule(-1, 5, 10)


def test_multi_gpu_request_exits_with_a_message(capsys, monkeypatch):
    monkeypatch.setattr("sys.argv", ["bench.py", "--gpus", "2", "--steps", "20", "--warmup", "5"])
    with pytest.raises(SystemExit) as e:
        bench.main()
    assert e.value.code == 2
    assert "one GPU" in capsys.readouterr().err
