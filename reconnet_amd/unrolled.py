"""Learned unrolled OS-SQS (no reference counterpart).

Each of the n unrolled iterations is one ordered-subsets SQS data pass
(ossqs.cu semantics, Siddon projector, all subsets once) followed by a learned
3x3 CNN prior with a residual connection (reconnet_amd.cnn.PriorCNN), which
replaces the TV surrogate of plain SQS-TV.  Every iteration has its own
network; they are trained greedily, stage by stage, on sinogram / image pairs
that the projector itself simulates from seeded phantoms
(reconnet_amd/train_unrolled.py).
"""

import os

import numpy as np
import torch

from .cnn import PriorCNN

__all__ = ["UnrolledOSSQS", "UnrolledGraph", "WEIGHTS_PATH", "load_nets", "save_nets"]

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unrolled_fan512.npz")


def save_nets(path, nets, meta=None):
    arrays = {}
    for k, net in enumerate(nets):
        arrays.update(net.to_numpy(prefix="it%d_" % k))
    if meta:
        for key, val in meta.items():
            arrays["meta_" + key] = np.asarray(val)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, **arrays)


def load_nets(path=WEIGHTS_PATH):
    z = np.load(path, allow_pickle=False)
    n = 0
    while ("it%d_w1" % n) in z.files:
        n += 1
    return [PriorCNN.from_numpy(z, prefix="it%d_" % k) for k in range(n)]


class UnrolledOSSQS:
    """Runs the unrolled reconstruction on an OsSqsTV engine that already holds
    the data (sinograms, weights, normaliser) of its slices."""

    def __init__(self, nets, lb=0.0):
        self.nets = nets
        self.lb = lb

    @property
    def niters(self):
        return len(self.nets)

    def reconstruct(self, eng, x0=None, chunk=None):
        g = eng.geo
        if x0 is None:
            x0 = torch.zeros((eng.nslices, g.nx, g.ny), dtype=torch.float32, device="cuda")
        eng.set_image(x0)
        x = None
        for net in self.nets:
            eng.run_pass(beta=0.0, lb=self.lb)
            x = eng.image(out=x)
            if chunk is None or chunk >= eng.nslices:
                x = net.forward(x)
            else:
                x = torch.cat([net.forward(x[i:i + chunk]) for i in range(0, eng.nslices, chunk)])
            eng.set_image(x)
        return x


class UnrolledGraph:
    """The whole unrolled reconstruction (image upload, every data pass and
    every CNN prior, image download) recorded once into a HIP graph and
    replayed per batch.

    The engine's sinograms, weights and normaliser are referenced by device
    address, so new slices of the same shape are swapped in with
    ``eng.set_data`` / ``eng.set_norm`` between replays.  ``x0`` is copied into
    the graph's own input buffer; the returned tensor is the graph's output
    buffer and is overwritten by the next replay."""

    def __init__(self, model, eng, x0=None):
        g = eng.geo
        self.x0 = torch.zeros((eng.nslices, g.nx, g.ny), dtype=torch.float32, device="cuda")
        if x0 is not None:
            self.x0.copy_(x0)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            model.reconstruct(eng, self.x0)       # warm-up outside the capture (workspaces, handles)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.out = model.reconstruct(eng, self.x0)
        torch.cuda.synchronize()

    def replay(self, x0=None):
        if x0 is not None:
            self.x0.copy_(x0)
        self.graph.replay()
        return self.out
