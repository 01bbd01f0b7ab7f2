"""Learned next-treelet prefetch: traces -> samples -> MLP -> prediction table.

No counterpart in the reference, whose RT unit prefetches by popularity only
(src/gpgpu-sim/shader.cc:3419-3559).  The pipeline uses the traversal's own
output: the per-ray node-access records of rtx_trace carry each node's
treelet (addrToTreeletID, vulkan_ray_tracing.cc:468-472), so a ray's treelet
sequence and its transitions come straight out of a traced batch.

Features of a transition cur -> next of a ray (rtx_mlp_sample):
  cur     the treelet the ray leaves;
  octant  the sign octant of the ray direction (bit0 x<0, bit1 y<0, bit2 z<0);
  depth   the tree level of cur's root / 32;
  hist    the last four treelets on the ray's path from the root to cur,
          nearest first.  A ray reaches a treelet only through its root, so
          these are cur's ancestor treelets, a function of cur.  This makes
          every feature except the octant a property of the treelet, and the
          network can be evaluated once per (treelet, octant) into a table
          that the traversal kernel reads with one 16-byte load per lane.
The target is the class of `next`: the n_classes-1 treelets entered most
often in the training traces, and one class for all others.

The table (rtx_set_prefetch_table) holds for each (treelet, octant) the
network's top-4 classes with their softmax probabilities; the "other" class
and treelets never left in the training traces get no entry, so those lanes
fall back to the heuristic.
"""
import numpy as np

from . import (MLP_SAMPLE_DTYPE, MODE_PLAIN, NO_TREELET, RAY_DTYPE, RECORD_DTYPE)

OCTANTS = 8
TX_DESCRIPTOR = 3


# ---- BVH structure: parent and tree level of every node ------------------------------
def node_tree(arena, tlas):
    """parent[slot] (NO_TREELET when none) and level[slot] for every node reachable from
    the TLAS, with the levels of tree_level_map (vulkan_ray_tracing.cc:2469, 2569, 2659):
    TLAS header 0, TLAS root 1, a child one deeper than its parent, a BLAS header and its
    root at the level of the instance leaf that opens them."""
    A = np.ascontiguousarray(arena, np.uint8)
    n_slots = len(A) // 64
    parent = np.full(n_slots, NO_TREELET, np.uint32)
    level = np.zeros(n_slots, np.uint32)
    u64 = lambda off: A[off[:, None] + np.arange(8)].copy().view(np.uint64)[:, 0]  # noqa: E731
    hs = tlas // 64
    root = np.array([tlas], np.uint64) + u64(np.array([tlas], np.int64))
    rs = int(root[0]) // 64
    parent[rs] = hs
    level[hs], level[rs] = 0, 1
    frontier = np.array([rs], np.int64)
    while len(frontier):
        base = frontier * 64
        child_off = A[base[:, None] + 12 + np.arange(4)].copy().view(np.int32)[:, 0].astype(np.int64)
        info = A[base[:, None] + 22 + np.arange(6)].astype(np.int64)
        size, ctype = info & 3, (info >> 2) & 15
        start = frontier + child_off
        cslot = start[:, None] + np.concatenate([np.zeros((len(frontier), 1), np.int64),
                                                 np.cumsum(size, axis=1)[:, :-1]], axis=1)
        valid = size > 0
        par = np.broadcast_to(frontier[:, None], cslot.shape)[valid]
        cs, ct = cslot[valid], ctype[valid]
        parent[cs] = par
        level[cs] = level[par] + 1
        nxt = [cs[ct == 0]]
        inst = cs[ct == 1]
        if len(inst):
            hdr = inst * 64 + u64(inst * 64 + 64).astype(np.int64)
            broot = hdr + u64(hdr).astype(np.int64)
            hslot, bslot = hdr // 64, broot // 64
            parent[hslot] = inst
            parent[bslot] = inst
            level[hslot] = level[inst]
            level[bslot] = level[inst]
            nxt.append(bslot)
        frontier = np.unique(np.concatenate(nxt))
    return parent, level


def treelet_context(roots, owner, parent, level):
    """Per treelet: the level of its root and its four nearest ancestor treelets."""
    rs = (np.asarray(roots, np.uint64) // 64).astype(np.int64)
    T = len(rs)
    par = parent[rs]
    ptree = np.full(T, NO_TREELET, np.uint32)
    ok = par != NO_TREELET
    ptree[ok] = owner[par[ok].astype(np.int64)]
    ptree[ptree == np.arange(T, dtype=np.uint32)] = NO_TREELET
    anc = np.full((T, 4), NO_TREELET, np.uint32)
    cur = np.arange(T, dtype=np.int64)
    alive = np.ones(T, bool)
    for h in range(4):
        p = np.where(alive, ptree[np.minimum(cur, T - 1)], NO_TREELET)
        anc[:, h] = p
        alive = p != NO_TREELET
        cur = np.where(alive, p, 0).astype(np.int64)
    return level[rs].astype(np.float32), anc


# ---- samples ---------------------------------------------------------------------------
def octants(rays):
    d = np.asarray(rays["dir"], np.float32)
    return ((d[:, 0] < 0).astype(np.uint32) | ((d[:, 1] < 0).astype(np.uint32) << 1) |
            ((d[:, 2] < 0).astype(np.uint32) << 2))


def transitions(rays, offsets, records):
    """(cur, next, ray) of every treelet change inside each ray's record stream
    (leaf-descriptor records, which lie inside the following leaf, skipped)."""
    counts = np.diff(np.asarray(offsets, np.int64))
    rid = np.repeat(np.arange(len(rays), dtype=np.int64), counts)
    keep = (records["type"] != TX_DESCRIPTOR) & (records["treelet"] != NO_TREELET)
    t, r = records["treelet"][keep], rid[keep]
    change = (t[1:] != t[:-1]) & (r[1:] == r[:-1])
    return t[:-1][change], t[1:][change], r[:-1][change]


def make_samples(cur, octant, root_level, anc, target):
    s = np.zeros(len(cur), MLP_SAMPLE_DTYPE)
    s["cur"] = cur
    s["octant"] = octant
    s["depth"] = root_level[cur] / 32.0
    s["hist"] = anc[cur]
    s["target"] = target
    return s


class LearnedPrefetcher:
    """Trains the next-treelet network on traces of a context's current scene and
    partition, and installs its prediction table."""

    def __init__(self, ctx, arena, tlas, n_classes=16384, seed=1, mode=MODE_PLAIN):
        self.ctx = ctx
        self.mode = mode
        self.seed = seed
        part = ctx.get_treelets()
        self.owner = part["owner"]
        self.T = len(part["roots"])
        parent, level = node_tree(arena, tlas)
        self.root_level, self.anc = treelet_context(part["roots"], self.owner, parent, level)
        self.n_classes = int(n_classes)
        self.class_treelet = None
        self.mlp = None
        self.seen = None

    def trace_transitions(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        _, off, rec = self.ctx.trace(rays, mode=self.mode, records=True)
        cur, nxt, rid = transitions(rays, off, rec)
        return cur, nxt, octants(rays)[rid]

    def _targets(self, nxt):
        cls = np.full(self.T, self.n_classes - 1, np.uint32)
        cls[self.class_treelet] = np.arange(len(self.class_treelet), dtype=np.uint32)
        return cls[nxt]

    def samples(self, cur, nxt, octs):
        return make_samples(cur, octs, self.root_level, self.anc, self._targets(nxt))

    def fit(self, rays, epochs=4, batch=1024, lr=2e-3, max_samples=4_000_000):
        """Trace `rays` with records, derive transitions, train the network."""
        cur, nxt, octs = self.trace_transitions(rays)
        vals, cnt = np.unique(nxt, return_counts=True)
        order = np.argsort(-cnt, kind="stable")
        self.class_treelet = vals[order[: self.n_classes - 1]].astype(np.uint32)
        self.seen = np.zeros(self.T, bool)
        self.seen[cur] = True
        s = self.samples(cur, nxt, octs)
        rng = np.random.default_rng(self.seed)
        s = s[rng.permutation(len(s))[:max_samples]]
        self.mlp = self.ctx.mlp(self.T, self.n_classes, self.seed)
        buf = self.mlp.upload(s)
        self.n_train = len(s)
        self.losses = []
        for _ in range(epochs):
            self.losses.append(self.mlp.train(buf, len(s), epochs=1, batch=batch, lr=lr))
        return self.losses

    def accuracy(self, rays):
        """Top-1 and top-4 accuracy of the network on the transitions of `rays`
        (targets outside the classes count as misses)."""
        cur, nxt, octs = self.trace_transitions(rays)
        s = self.samples(cur, nxt, octs)
        top, _ = self.mlp.predict(s)
        cls = top.astype(np.int64)
        pred_t = np.where(cls < len(self.class_treelet), self.class_treelet[np.minimum(cls, len(self.class_treelet) - 1)],
                          NO_TREELET)
        hit = pred_t == nxt[:, None]
        return float(hit[:, 0].mean()), float(hit.any(axis=1).mean()), len(s)

    def table(self):
        """Packed prediction table, uint32 [T * 8, 4]."""
        t_ids = np.nonzero(self.seen)[0].astype(np.uint32)
        cur = np.repeat(t_ids, OCTANTS)
        octs = np.tile(np.arange(OCTANTS, dtype=np.uint32), len(t_ids))
        s = make_samples(cur, octs, self.root_level, self.anc, np.zeros(len(cur), np.uint32))
        top, prob = self.mlp.predict(s)
        C1 = len(self.class_treelet)
        ok = top < C1
        tre = np.where(ok, self.class_treelet[np.minimum(top, max(C1 - 1, 0)).astype(np.int64)], 0)
        q = np.clip(np.rint(prob * 255.0), 0, 255).astype(np.uint32)
        ent = np.where(ok & (q > 0), (tre + 1) | (q << 24), 0).astype(np.uint32)
        out = np.zeros((self.T * OCTANTS, 4), np.uint32)
        out[cur.astype(np.int64) * OCTANTS + octs] = ent
        return out

    def install(self):
        tab = self.table()
        self.ctx.set_prefetch_table(tab)
        return tab

    def close(self):
        if self.mlp is not None:
            self.mlp.close()
            self.mlp = None


# ---- one network over several scenes -----------------------------------------------------
def scene_transitions(ctx, arena, tlas, budget, rays, mode=MODE_PLAIN):
    """Set the scene and its partition on ctx, trace `rays` with records and return the
    scene's treelet context and its transitions (cur, next, octant)."""
    ctx.set_scene(arena, tlas)
    ctx.build_treelets(budget)
    part = ctx.get_treelets()
    parent, level = node_tree(arena, tlas)
    root_level, anc = treelet_context(part["roots"], part["owner"], parent, level)
    rays = np.ascontiguousarray(rays, RAY_DTYPE)
    _, off, rec = ctx.trace(rays, mode=mode, records=True)
    cur, nxt, rid = transitions(rays, off, rec)
    return {"T": len(part["roots"]), "root_level": root_level, "anc": anc, "cur": cur, "nxt": nxt,
            "oct": octants(rays)[rid]}


class JointPrefetcher:
    """One next-treelet network over the treelets of several scenes: scene k's treelet
    t is embedding row offset[k] + t, the trunk is shared, and the classes are the
    n_classes - 1 next treelets (of any scene) seen most often.  `sources` are
    scene_transitions() results."""

    def __init__(self, ctx, sources, n_classes=16384, seed=1):
        self.ctx, self.seed, self.n_classes = ctx, seed, int(n_classes)
        self.offset = np.cumsum([0] + [s["T"] for s in sources])[:-1].astype(np.int64)
        self.T = int(sum(s["T"] for s in sources))
        self.sizes = [s["T"] for s in sources]
        self.root_level = np.concatenate([s["root_level"] for s in sources]).astype(np.float32)
        ancs = []
        for s, o in zip(sources, self.offset):
            a = s["anc"].astype(np.int64)
            ancs.append(np.where(s["anc"] == NO_TREELET, NO_TREELET, a + o).astype(np.uint32))
        self.anc = np.concatenate(ancs)
        self.cur = np.concatenate([s["cur"].astype(np.int64) + o for s, o in zip(sources, self.offset)])
        self.nxt = np.concatenate([s["nxt"].astype(np.int64) + o for s, o in zip(sources, self.offset)])
        self.oct = np.concatenate([s["oct"] for s in sources]).astype(np.uint32)
        vals, cnt = np.unique(self.nxt, return_counts=True)
        order = np.argsort(-cnt, kind="stable")
        self.class_treelet = vals[order[: self.n_classes - 1]].astype(np.int64)
        self.seen = np.zeros(self.T, bool)
        self.seen[self.cur] = True
        self.mlp = None

    def _samples(self, cur, octs, nxt):
        cls = np.full(self.T, self.n_classes - 1, np.uint32)
        cls[self.class_treelet] = np.arange(len(self.class_treelet), dtype=np.uint32)
        return make_samples(cur.astype(np.uint32), octs, self.root_level, self.anc, cls[nxt])

    def fit(self, epochs=2, batch=1024, lr=2e-3, max_samples=3_000_000):
        s = self._samples(self.cur, self.oct, self.nxt)
        rng = np.random.default_rng(self.seed)
        s = s[rng.permutation(len(s))[:max_samples]]
        self.mlp = self.ctx.mlp(self.T, self.n_classes, self.seed)
        buf = self.mlp.upload(s)
        self.n_train = len(s)
        self.losses = [self.mlp.train(buf, len(s), epochs=1, batch=batch, lr=lr) for _ in range(epochs)]
        return self.losses

    def accuracy(self, k, cur, nxt, octs):
        """Top-1 / top-4 accuracy on transitions (scene-local ids) of scene k."""
        o = self.offset[k]
        s = self._samples(cur.astype(np.int64) + o, octs.astype(np.uint32), nxt.astype(np.int64) + o)
        top, _ = self.mlp.predict(s)
        C1 = len(self.class_treelet)
        pred = np.where(top < C1, self.class_treelet[np.minimum(top, C1 - 1).astype(np.int64)], -1)
        hit = pred == (nxt.astype(np.int64) + o)[:, None]
        return float(hit[:, 0].mean()), float(hit.any(axis=1).mean())

    def table(self, k):
        """Prediction table of scene k (its treelets and its classes only), uint32 [T_k * 8, 4]."""
        o, n = int(self.offset[k]), self.sizes[k]
        t_ids = np.nonzero(self.seen[o:o + n])[0].astype(np.int64)
        cur = np.repeat(t_ids + o, OCTANTS)
        octs = np.tile(np.arange(OCTANTS, dtype=np.uint32), len(t_ids))
        s = make_samples(cur.astype(np.uint32), octs, self.root_level, self.anc, np.zeros(len(cur), np.uint32))
        top, prob = self.mlp.predict(s)
        C1 = len(self.class_treelet)
        tre = np.where(top < C1, self.class_treelet[np.minimum(top, C1 - 1).astype(np.int64)], -1)
        local = tre - o
        ok = (tre >= 0) & (local >= 0) & (local < n)
        q = np.clip(np.rint(prob * 255.0), 0, 255).astype(np.uint32)
        ent = np.where(ok & (q > 0), (np.where(ok, local, 0).astype(np.uint32) + 1) | (q << 24), 0).astype(np.uint32)
        out = np.zeros((n * OCTANTS, 4), np.uint32)
        out[(cur - o) * OCTANTS + octs] = ent
        return out

    def close(self):
        if self.mlp is not None:
            self.mlp.close()
            self.mlp = None
