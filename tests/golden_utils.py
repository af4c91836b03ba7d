"""Load the committed golden fixtures (reference outputs) as designs + arrays."""
import os

import numpy as np

from eplace_amd.design import Design

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


class Golden:
    def __init__(self, name):
        z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
        self.a = {k: z[k] for k in z.files}
        self.s = dict(zip(self.a["scalar_names"].tolist(), self.a["scalar_values"].tolist()))
        self.name = name

    def __getitem__(self, k):
        return self.a[k]

    def design(self) -> Design:
        a = self.a
        n = a["nodes_whxy"]
        t = a["terms_whxyll"] if "terms_whxyll" in a else np.zeros((0, 6), np.float32)
        rows = a["rows"]
        N = len(n)
        return Design(
            name=self.name, node_names=[f"o{i}" for i in range(N)],
            node_w=n[:, 0].copy(), node_h=n[:, 1].copy(), node_x=n[:, 2].copy(), node_y=n[:, 3].copy(),
            term_names=[f"t{i}" for i in range(len(t))],
            term_w=t[:, 0].copy(), term_h=t[:, 1].copy(), term_x=t[:, 2].copy(), term_y=t[:, 3].copy(),
            term_ni=a["terms_ni"].astype(np.uint8),
            net_start=a["net_start"].astype(np.int64), pin_obj=a["pin_obj"].astype(np.int32),
            pin_dx=a["pin_offset"][:, 0].copy(), pin_dy=a["pin_offset"][:, 1].copy(),
            row_bottom=rows[:, 0].copy(), row_height=rows[:, 1].copy(), row_x0=rows[:, 2].copy(),
            row_x1=rows[:, 3].copy(), site_width=float(rows[0, 4]),
        )

    @property
    def target_density(self):
        if "target_density" in self.a:
            return float(self.a["target_density"])
        return 0.8


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))
