"""bench.py's command-line contract, checked without a GPU: one GPU only (a
request for more exits with a message before any device work), and every
--precision choice maps to a library precision with its roofline peak, its
dtype label and the element type of the probed transform kernel."""
import os
import subprocess
import sys

import bench
import se_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, r
    assert "one GPU" in r.stderr and r.stdout == ""


def test_every_precision_choice_is_complete():
    old = sys.argv
    try:
        sys.argv = ["bench.py"]
        a = bench.parse()
    finally:
        sys.argv = old
    assert a.gpus == 1 and a.steps >= 1 and a.warmup >= 0
    src = open(os.path.join(ROOT, "bench.py")).read()
    start = src.index('"--precision", default=None, choices=[') + len('"--precision", default=None, choices=[')
    parser_choices = [c.strip().strip('"') for c in src[start:src.index("]", start)].split(",")]
    codes = set()
    for name in parser_choices:
        code = bench.prec_code(name)
        assert code in (se_amd.PREC_FP64, se_amd.PREC_FP32_IR, se_amd.PREC_BF16_IR, se_amd.PREC_FP64_FP32C,
                        se_amd.PREC_FP64_BF16C)
        codes.add(code)
        assert name in bench.MFMA_PEAK_TFLOPS and name in bench.DTYPE and name in bench.GEMM_T
    assert len(codes) == len(parser_choices)
