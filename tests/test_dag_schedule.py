"""The batched solve's task scheduler (host code, no GPU).

The persistent task kernel takes tasks in list order and waits on counters
of earlier tasks; an order in which some wait refers to a later task would
hang it.  These tests build dag_schedule.h into a small host program and
check, for random batches and worker counts, that the list schedule is a
complete topological order under the kernel's exact wait rules, and that the
checker the plan runs before every launch rejects corrupted orders (a
diagonal task ahead of the column it waits for, two tasks of a TRTRI column
swapped, a task missing).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cuda-cholesky_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    out = str(tmp_path_factory.mktemp("dag") / "dag_schedule_check")
    src = os.path.join(ROOT, "tests", "dag_schedule_check.cpp")
    subprocess.run([cxx, "-O2", "-std=c++17", "-I", CSRC, src, "-o", out], check=True)
    return out


def expected_tasks(sizes):
    total = 0
    for n in sizes:
        T = (n + 63) // 64
        total += (T + T * (T - 1) // 2 + T * (T - 1) // 2 + T * (T + 1) // 2 + 1) if T else 1
    return total


@pytest.mark.parametrize("seed,count,nmax,workers", [
    (1, 1, 64, 1), (2, 1, 2048, 1), (3, 3, 300, 2), (4, 40, 700, 7),
    (5, 200, 2048, 512), (6, 512, 2048, 512), (7, 64, 130, 1000), (8, 20, 4096, 3)])
def test_schedule_is_a_valid_order(checker, seed, count, nmax, workers):
    r = subprocess.run([checker, str(seed), str(count), str(nmax), str(workers)], check=True,
                       capture_output=True, text=True)
    ntasks, check, mut1, mut2, mut3 = map(int, r.stdout.split())
    assert check == -1, r.stdout
    assert ntasks > 0
    # corrupted orders are rejected (a wait on a later task, or a task missing)
    if nmax > 64:
        assert mut1 != -1 and mut2 != -1, r.stdout
    assert mut3 != -1, r.stdout


def test_task_count_matches_formula(checker):
    """One matrix of order n (the harness's first matrix has order nmax)."""
    for n in (1, 63, 64, 65, 640, 2048):
        r = subprocess.run([checker, "1", "1", str(n), "4"], check=True, capture_output=True, text=True)
        assert int(r.stdout.split()[0]) == expected_tasks([n]), (n, r.stdout)
