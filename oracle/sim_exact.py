"""ORACLE -- test infrastructure only. The exact distribution of the dye
tracks whatprot's simulator draws for one dye-seq (generate_dye_track,
simulation/generate-dye-track.cc:29-110), by a forward pass over the
simulator's own state instead of sampling. Small dye-seqs only: the state
holds which dyes are still alive.

The simulator's process, step for step:
  - duds: each dye is removed with its channel's p_dud (:48-54);
  - initial block with p_initial_block (:63-66);
  - for t < T, while the Edman position e is inside the dye-seq: record the
    counts of live dyes at positions >= e (:69-72); cyclic block with
    p_cyclic_block (:74-76); detach with p_detach[t], after which every record
    is zero (:78-81); unless blocked, Edman fails with p_edman_failure or else
    removes position e and advances (:83-90); every live dye at positions >= e
    bleaches with its channel's p_bleach (:92-99);
  - once e reaches the end, the record is zero (:101-105).
Trailing '.' are not part of the dye-seq (common/dye-seq.cc:23-40).
"""
from collections import defaultdict
from itertools import combinations


def track_distribution(model, T, dye_string, detach_offset=0):
    """{track (T * C counts, t-major): probability} of generate_dye_track.

    detach_offset=1 reads the detach rate of cycle t from p_detach[t + 1], an
    off-by-one the tests plant to show that a faithful sampler rejects it.
    (Swapping Edman and Bleach within a cycle would not do: the dye at the
    Edman position is gone after a successful cycle either way, so the
    tracks have the same distribution.)"""
    C = model.num_channels
    seq = [int(ch) if ch.isdigit() else -1 for ch in dye_string.rstrip(".")]
    L = len(seq)
    dyes = [i for i, c in enumerate(seq) if c >= 0]
    ch = [seq[i] for i in dyes]
    cm = model.channel_models
    pib, pcb, pf = model.p_initial_block, model.p_cyclic_block, model.p_edman_failure

    def bleach(mask, e):
        """(mask after bleaching the live dyes at positions >= e, probability)."""
        live = [j for j, i in enumerate(dyes) if (mask >> j) & 1 and i >= e]
        out = []
        for k in range(len(live) + 1):
            for gone in combinations(live, k):
                p = 1.0
                for j in live:
                    q = cm[ch[j]].p_bleach
                    p *= q if j in gone else 1.0 - q
                m2 = mask
                for j in gone:
                    m2 &= ~(1 << j)
                out.append((m2, p))
        return out

    states = defaultdict(float)  # (e, blocked, live mask, track so far) -> probability
    for mask in range(1 << len(dyes)):
        p = 1.0
        for j in range(len(dyes)):
            q = cm[ch[j]].p_dud
            p *= 1.0 - q if (mask >> j) & 1 else q
        states[(0, True, mask, ())] += p * pib
        states[(0, False, mask, ())] += p * (1.0 - pib)
    for t in range(T):
        pd = model.p_detach_at(t + detach_offset)
        new = defaultdict(float)
        for (e, blocked, mask, tr), p in states.items():
            if e >= L:
                new[(e, blocked, mask, tr + (0,) * C)] += p
                continue
            counts = [0] * C
            for j, i in enumerate(dyes):
                if (mask >> j) & 1 and i >= e:
                    counts[ch[j]] += 1
            tr2 = tr + tuple(counts)
            for cyc, pc in ((True, pcb), (False, 1.0 - pcb)):
                blk = blocked or cyc
                new[(L, blk, mask, tr2)] += p * pc * pd
                pn = p * pc * (1.0 - pd)
                moves = [(e, pn)] if blk else [(e + 1, pn * (1.0 - pf)), (e, pn * pf)]
                for e2, q in moves:
                    for m2, pb in bleach(mask, e2):
                        new[(e2, blk, m2, tr2)] += q * pb
        states = new
    dist = defaultdict(float)
    for (_, _, _, tr), p in states.items():
        dist[tr] += p
    return dict(dist)
