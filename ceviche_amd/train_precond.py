"""Trains the learned FDFD preconditioner on the GPU and saves its weights.

    python -m ceviche_amd.train_precond [--size 256] [--structures 60]
        [--steps 30000] [--batch 16] [--out ceviche_amd/data/unet_fdfd_ez.npz]

Pairs come from plain BiCGSTAB runs on seeded binary-blob structures
(workloads.fdfd_learned_blobs, training seeds 0 .. structures-1); training
uses the network's own forward / backward kernels (learned.train).  The
saved file holds W, b, levels, nf, the output gain and the training record.
"""
import argparse
import os
import sys
import time

import numpy as np
import torch

from . import learned

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'data', 'unet_fdfd_ez.npz')


def log(*a):
    print(*a, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--size', type=int, default=256)
    ap.add_argument('--structures', type=int, default=60)
    ap.add_argument('--snaps', type=int, default=24)
    ap.add_argument('--steps', type=int, default=30000)
    ap.add_argument('--batch', type=int, default=16)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--levels', type=int, default=6)
    ap.add_argument('--w-res', type=float, default=1.0)
    ap.add_argument('--w-err', type=float, default=0.1)
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--out', default=DEFAULT_OUT)
    a = ap.parse_args(argv)
    t0 = time.time()
    pairs, stats = learned.generate_pairs(a.structures, N=a.size, snaps=a.snaps, chunk=32, seed0=0, log=log)
    torch.cuda.synchronize()
    t_pairs = time.time() - t0
    log('pairs %d from %d structures in %.1f s (mean BiCGSTAB iterations %.0f, gain %.3f)'
        % (pairs.n, a.structures, t_pairs, float(np.mean(stats['iterations'])), pairs.gain))
    net = learned.UNet(levels=a.levels, seed=a.seed)
    before = learned.evaluate(net, pairs)
    t0 = time.time()
    hist = learned.train(net, pairs, steps=a.steps, batch=a.batch, lr=a.lr, seed=a.seed, log=log,
                         log_every=1000, w_res=a.w_res, w_err=a.w_err)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    after = learned.evaluate(net, pairs)
    log('training %.1f s; residual ratio %.4f -> %.4f, error ratio %.4f -> %.4f'
        % (t_train, before[0], after[0], before[1], after[1]))
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out, gain=np.float64(pairs.gain), train_size=np.int64(a.size), train_pairs=np.int64(pairs.n),
             train_steps=np.int64(a.steps), train_seconds=np.float64(t_train), pair_seconds=np.float64(t_pairs),
             residual_ratio=np.float64(after[0]), error_ratio=np.float64(after[1]),
             loss_history=np.array(hist, dtype=np.float64))
    log('saved', a.out)
    return 0


if __name__ == '__main__':
    sys.exit(main())
