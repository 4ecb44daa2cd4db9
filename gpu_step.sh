#!/bin/bash
# temporary driver script (deleted after the call)
cd "$GRAFT_REPO_ROOT" || exit 1
export TMPDIR=/tmp
R=$GRAFT_REPO_ROOT
echo "step: tuner + mixed tests"
timeout -k 10 400 python -u -m pytest tests/test_gpu_tuner.py tests/test_gpu_mixed.py -x -v -s --timeout 200 --timeout-method thread > $R/gpurun_out/t11a.log 2>&1 || { tail -40 $R/gpurun_out/t11a.log; exit 1; }
grep -E "componentwise|passed|failed" $R/gpurun_out/t11a.log | tail -24
echo "step: gpu tests"
timeout -k 10 900 python -u -m pytest tests -m gpu -x -q --timeout 300 --timeout-method thread > $R/gpurun_out/t11.log 2>&1 || { tail -40 $R/gpurun_out/t11.log; exit 1; }
tail -3 $R/gpurun_out/t11.log
echo "step: bench ozaki-I"
timeout -k 10 300 python bench.py --workload gemm6k_f64 --precision ozaki_int8 --steps 3 --warmup 6 --no-cpu > $R/gpurun_out/b_oz.json 2> $R/gpurun_out/b_oz.err || { tail -20 $R/gpurun_out/b_oz.err; exit 1; }
python3 -c "import json; d=json.load(open('$R/gpurun_out/b_oz.json')); print(d['ms_per_step'], d['roofline'], d['errors_vs_fp64_oracle'])"
echo "step: done"
