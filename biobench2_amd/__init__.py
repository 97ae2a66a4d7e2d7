"""MI355X (gfx950) HMMER3 hmmsearch filter pipeline: SSV/MSV, Viterbi and
Forward/Backward stages behind the p7_gpu_* C ABI (include/p7_gpu.h).

The hot path is libp7gpu.so (host C + HIP kernels); p7gpu is its ctypes
binding, seqio turns text into digital residues and synth generates the
synthetic workloads.
"""
