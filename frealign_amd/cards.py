"""The card deck of a Frealign run (cards 1-18, read from standard input;
card1.f ... cards15to18.f), as the heredocs of mult_refine_n.com and
mult_reconstruct_n.com write it.

Numeric cards are read as Fortran list-directed input reads them: values
separated by commas or blanks, T/F for logicals, the rest of the record
(comments such as '!RO,RI,...') ignored once the card's items are read.
Filename cards are the first 200 characters of their line, trimmed.  Cards
6-12 repeat per dataset until a card 6 with RELMAG = 0 (or negative: no
reconstruction, card6.f:47-50).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field


class CardError(ValueError):
    pass


def _values(line: str, n: int, types, card: str):
    """The first n list-directed items of a record."""
    toks = []
    cur = ""
    for ch in line:
        if ch in ", \t":
            if cur:
                toks.append(cur)
                cur = ""
        else:
            cur += ch
        if len(toks) >= n:
            break
    if cur and len(toks) < n:
        toks.append(cur)
    if len(toks) < n:
        raise CardError(f"card {card}: {len(toks)} values, {n} expected: {line!r}")
    out = []
    for t, ty in zip(toks[:n], types):
        try:
            if ty is bool:
                u = t.strip(".").upper()
                if u[:1] not in ("T", "F"):
                    raise ValueError(t)
                out.append(u[:1] == "T")
            elif ty is int:
                out.append(int(float(t)) if "." in t else int(t))
            elif ty is str:
                out.append(t)
            else:
                out.append(float(t.replace("d", "e").replace("D", "e")))
        except ValueError as e:
            raise CardError(f"card {card}: bad value {t!r}: {line!r}") from e
    return out


@dataclass
class Dataset:
    relmag: float
    dstep: float
    target: float
    thresh: float
    cs: float          # mm as on the card
    akv: float         # kV as on the card
    tx: float
    ty: float
    rrec: float        # A as on card 7
    rmax1: float
    rmax2: float
    rclas: float
    dfstd: float
    rbfact: float
    finpat1: str
    finpat2: str
    finpar: str
    foutpar: str
    foutsh: str

    @property
    def wl(self) -> float:
        """card6.f:51-52: the electron wavelength (A) at AKV kV."""
        akv = 1000.0 * self.akv
        return 12.26 / math.sqrt(akv + 0.9785 * akv ** 2 / (10.0 ** 6.0))


@dataclass
class Deck:
    cform: str
    iflag: int
    fmag: bool
    fdef: bool
    fastig: bool
    fpart: bool
    iewald: int
    fbeaut: bool
    ffilt: bool
    fbfact: bool
    fmatch: bool
    ifsc: int
    fdump: bool
    imem: int
    interp: int
    ro: float
    ri: float
    psize: float
    mw: float
    wgh: float
    xstd: float
    pbc: float
    boff: float
    dang: float
    itmax: int
    ipmax: int
    pmask: list        # as on the card: PSI, THETA, PHI, SHX, SHY
    dmask: list
    ifirst: int
    ilast: int
    asym: str
    datasets: list = field(default_factory=list)
    reconstruct: bool = True
    f3d: str = ""
    fweigh: str = ""
    map1: str = ""
    map2: str = ""
    fpha: str = ""
    fpoi: str = ""

    @property
    def pmask_internal(self):
        """card3.f:47-50: PSI and PHI swap places inside the program."""
        m = list(self.pmask)
        m[0], m[2] = m[2], m[0]
        return m

    def point_group(self):
        """(letter, n) of card 5's ASYM (C1, Cn, Dn, T, O, I)."""
        a = self.asym.strip().upper()
        if not a or a[0] not in "CDTOI":
            raise CardError(f"card 5: symmetry {self.asym!r} not supported here")
        n = int(a[1:]) if a[1:].isdigit() else (1 if a[0] == "C" else 0)
        return a[0], n


def parse_deck(text: str) -> Deck:
    lines = [ln.rstrip("\n") for ln in text.splitlines()]
    pos = 0

    def nxt(card):
        nonlocal pos
        while pos < len(lines) and not lines[pos].strip():
            pos += 1
        if pos >= len(lines):
            raise CardError(f"card {card}: end of input")
        ln = lines[pos]
        pos += 1
        return ln

    # card 1 (card1.f:60-63): CFORM in column 1, the flags from column 3
    c1 = nxt("1")
    cform = c1[:1].upper()
    v = _values(c1[2:], 14, [int, bool, bool, bool, bool, int, bool, bool, bool, bool, int, bool,
                             int, int], "1")
    d = dict(cform=cform, iflag=v[0], fmag=v[1], fdef=v[2], fastig=v[3], fpart=v[4], iewald=v[5],
             fbeaut=v[6], ffilt=v[7], fbfact=v[8], fmatch=v[9], ifsc=v[10], fdump=v[11],
             imem=v[12], interp=v[13])
    if d["iflag"] == 0:  # card1.f:99-117
        d["fmag"] = d["fdef"] = d["fastig"] = False
    # card 2 (card2.f:22-24)
    v = _values(nxt("2"), 11, [float] * 9 + [int, int], "2")
    if v[0] == 0.0:
        raise CardError("card 2: RO = 0")
    d.update(ro=v[0], ri=v[1], psize=v[2], mw=v[3], wgh=v[4], xstd=v[5], pbc=v[6], boff=v[7],
             dang=v[8], itmax=v[9], ipmax=max(v[10], 1))
    # card 3 (card3.f:24-44): PMASK (5) and, optionally, DMASK (4)
    c3 = nxt("3")
    try:
        v = _values(c3, 9, [int] * 5 + [float] * 4, "3")
        pmask, dmask = v[:5], v[5:]
    except CardError:
        pmask, dmask = _values(c3, 5, [int] * 5, "3"), [-1.0, 0.0, 0.0, 0.0]
    if any(m not in (0, 1) for m in pmask):
        raise CardError("card 3: invalid PMASK values")
    d.update(pmask=pmask, dmask=dmask)
    v = _values(nxt("4"), 2, [int, int], "4")
    d.update(ifirst=v[0], ilast=v[1])
    d["asym"] = nxt("5")[:3].strip()
    deck = Deck(**d)
    # cards 6-12 per dataset, until RELMAG <= 0 (card6.f:31-50)
    while True:
        c6 = nxt("6")
        try:
            v6 = _values(c6, 8, [float] * 8, "6")
        except CardError:
            v6 = _values(c6, 6, [float] * 6, "6") + [0.0, 0.0]
        if v6[0] <= 0.0:
            deck.reconstruct = v6[0] == 0.0
            break
        c7 = nxt("7")
        try:
            v7 = _values(c7, 6, [float] * 6, "7")
        except CardError:
            r = _values(c7, 4, [float] * 4, "7")
            v7 = [r[0], r[1], r[2], r[2], 100.0, r[3]]
        v7[4] = abs(v7[4])  # card7.f:40: DFSTD = ABS(DFSTD)
        fin1 = nxt("8").strip()[:200]
        fin2 = nxt("9").strip()[:200]
        if deck.iflag >= 0:
            finpar = nxt("10").strip()[:200]
        else:
            raise CardError("card 10b (creating a parameter file, IFLAG < 0) is not supported")
        fout = nxt("11").strip()[:200]
        fsh = nxt("12").strip()[:200]
        deck.datasets.append(Dataset(*v6, *v7, fin1, fin2, finpar, fout, fsh))
    for name in ("f3d", "fweigh", "map1", "map2", "fpha", "fpoi"):
        if pos < len(lines) and lines[pos].strip():
            setattr(deck, name, nxt(name).strip()[:200])
    if not deck.datasets:
        raise CardError("no dataset (cards 6-12)")
    return deck
