"""On-disk formats around the path: Stokes images and movies as FITS files in ehtim's layout, and interferometric
observations from uvfits files (the reference's src/Fit/fits.jl:1-107 and src/Fit/uvfits.jl:1-205), read and
written with numpy (no FITS library is assumed).

    img, hdr = read_stokes_fits(path)                       # (nx, ny, 4) Jy/pixel, header quantities
    movie, camera, L = read_stokes_movie(paths, M_solar=..., D_pc=..., sigma=...)
    write_stokes_fits(path, image_cgs, psize_M, M_solar=..., D_pc=..., freq=..., mjd=...)
    obs = read_uvfits(path)                                 # Observation: rows of baselines, Stokes visibilities

FITS basics used here (FITS Standard 4.0): 2880-byte blocks of 80-character header cards ending at END, big-endian
data after the header padded to whole blocks; images are stored with the first axis fastest; a random-groups
primary HDU (GROUPS = T, NAXIS1 = 0) holds GCOUNT groups of PCOUNT parameters followed by the array of the
remaining axes; binary tables (XTENSION = 'BINTABLE') hold NAXIS2 rows of NAXIS1 bytes with TFORMn fields.
"""
import math

import numpy as np

from .api import Camera, gravitational_radius

BLOCK = 2880
PC = 3.085678e18
JY = 1e-23
CL = 2.99792458e10


# ---- the FITS container ---------------------------------------------------------------------------------------
def _parse_value(s):
    s = s.strip()
    if not s:
        return None
    if s.startswith("'"):
        end = s.find("'", 1)
        while end != -1 and end + 1 < len(s) and s[end + 1] == "'":
            end = s.find("'", end + 2)
        return s[1:end].replace("''", "'").rstrip()
    v = s.split("/", 1)[0].strip()
    if v == "T":
        return True
    if v == "F":
        return False
    try:
        return int(v)
    except ValueError:
        pass
    try:
        return float(v.replace("D", "E"))
    except ValueError:
        return v


def _read_header(buf, pos):
    """Header cards from byte pos: (dict, list of keys in order, position of the data)."""
    h, keys = {}, []
    while True:
        block = buf[pos:pos + BLOCK]
        if len(block) < BLOCK:
            raise ValueError("truncated FITS header")
        pos += BLOCK
        for i in range(0, BLOCK, 80):
            card = block[i:i + 80].decode("ascii", "replace")
            key = card[:8].strip()
            if key == "END":
                return h, keys, pos
            if card[8:10] == "= ":
                h[key] = _parse_value(card[10:])
                keys.append(key)


def _data_bytes(h):
    bitpix = abs(int(h["BITPIX"])) // 8
    naxis = int(h.get("NAXIS", 0))
    if naxis == 0:
        return 0
    dims = [int(h[f"NAXIS{i}"]) for i in range(1, naxis + 1)]
    if h.get("GROUPS") is True and dims[0] == 0:
        n = int(np.prod(dims[1:])) + int(h.get("PCOUNT", 0))
        return bitpix * int(h.get("GCOUNT", 1)) * n
    return bitpix * (int(np.prod(dims)) + int(h.get("PCOUNT", 0))) * int(h.get("GCOUNT", 1))


def _dtype(bitpix):
    return {8: ">u1", 16: ">i2", 32: ">i4", 64: ">i8", -32: ">f4", -64: ">f8"}[int(bitpix)]


def read_hdus(path):
    """Every HDU of a FITS file as (header dict, raw data bytes)."""
    buf = open(path, "rb").read()
    pos, out = 0, []
    while pos < len(buf):
        if not buf[pos:pos + 80].strip(b" \x00"):
            break
        h, _, pos = _read_header(buf, pos)
        n = _data_bytes(h)
        out.append((h, buf[pos:pos + n]))
        pos += (n + BLOCK - 1) // BLOCK * BLOCK
    return out


def _image(h, data):
    naxis = int(h["NAXIS"])
    dims = [int(h[f"NAXIS{i}"]) for i in range(1, naxis + 1)]
    a = np.frombuffer(data, dtype=_dtype(h["BITPIX"]), count=int(np.prod(dims))).astype(np.float64)
    a = a * float(h.get("BSCALE", 1.0)) + float(h.get("BZERO", 0.0))
    return a.reshape(dims[::-1]).T                  # first axis fastest: index (x, y, ...)


def _card(key, value, comment=""):
    if isinstance(value, bool):
        v = f"{'T' if value else 'F':>20}"
    elif isinstance(value, (int, np.integer)):
        v = f"{int(value):>20d}"
    elif isinstance(value, (float, np.floating)):
        v = f"{repr(float(value)).upper().replace('INF', '1E999'):>20}"
    else:
        s = str(value).replace("'", "''")
        v = f"'{s:<8}'"
    c = f"{key:<8}= {v}"
    if comment:
        c += f" / {comment}"
    return c[:80].ljust(80)


def _write_hdu(f, cards, data):
    head = "".join(cards) + "END".ljust(80)
    head = head.ljust((len(head) + BLOCK - 1) // BLOCK * BLOCK)
    f.write(head.encode("ascii"))
    raw = data.tobytes()
    f.write(raw + b"\x00" * ((-len(raw)) % BLOCK))


# ---- Stokes images and movies (fits.jl) ---------------------------------------------------------------------------
def time_unit(M_solar):
    """GM/c^3 [s] for a mass in solar masses (fits.jl:7)."""
    return gravitational_radius(M_solar) / CL


def read_stokes_fits(path):
    """read_stokes_fits(path) (fits.jl:9-40): the HDUs whose STOKES is I, Q, U, V (a missing one counts as zero) as
    an (nx, ny, 4) array in the file's units indexed (x, y), and the header quantities psize_deg (|CDELT2|), freq
    (FREQ, Hz), mjd, bunit. The first image HDU without STOKES counts as I."""
    imgs, hdr = {}, None
    for h, data in read_hdus(path):
        if int(h.get("NAXIS", 0)) < 2 or h.get("GROUPS") is True:
            continue
        if h.get("XTENSION") not in (None, "IMAGE"):
            continue
        if hdr is None:
            hdr = h
        st = str(h.get("STOKES", "I" if not imgs else " ")).strip().upper()[:1]
        if st in "IQUV" and st:
            imgs[st] = _image(h, data)[:, :, 0, 0] if int(h["NAXIS"]) > 2 else _image(h, data)
    if "I" not in imgs:
        raise ValueError(f"no Stokes I image in {path}")
    n = imgs["I"].shape
    S = np.stack([imgs.get(c, np.zeros(n)) for c in "IQUV"], axis=-1)
    header = {"psize_deg": abs(float(hdr.get("CDELT2", hdr.get("CDELT1")))),
              "freq": float(hdr["FREQ"]) if "FREQ" in hdr else math.nan,
              "mjd": float(hdr["MJD"]) if "MJD" in hdr else math.nan,
              "bunit": str(hdr.get("BUNIT", "JY/PIXEL")).strip().upper()}
    return S, header


def write_stokes_fits(path, image, psize_M, M_solar, D_pc, freq, mjd, source="KerrSplat"):
    """write_stokes_fits(path, image, psize_M; M_solar, D_pc, freq, mjd, source) (fits.jl:83-105): a Stokes image
    (nx, ny, 4) in cgs intensity, indexed (x, y) with x toward west = +alpha and y north, as an ehtim-style FITS file
    in Jy/pixel with one HDU per Stokes parameter (the primary holds I)."""
    img = np.asarray(image, dtype=np.float64)
    L = gravitational_radius(M_solar)
    psize_rad = psize_M * L / (D_pc * PC)
    omega = psize_rad ** 2
    nx, ny = img.shape[:2]
    with open(path, "wb") as f:
        for k, st in enumerate("IQUV"):
            cards = ([_card("SIMPLE", True), ] if k == 0 else [_card("XTENSION", "IMAGE")])
            cards += [_card("BITPIX", -64), _card("NAXIS", 2), _card("NAXIS1", nx), _card("NAXIS2", ny)]
            cards += [_card("EXTEND", True)] if k == 0 else [_card("PCOUNT", 0), _card("GCOUNT", 1)]
            cards += [_card("OBJECT", source), _card("CTYPE1", "RA---SIN"), _card("CTYPE2", "DEC--SIN"),
                      _card("CDELT1", -psize_rad * 180 / math.pi), _card("CDELT2", psize_rad * 180 / math.pi),
                      _card("CRPIX1", (nx + 1) / 2), _card("CRPIX2", (ny + 1) / 2), _card("CRVAL1", 0.0), _card("CRVAL2", 0.0),
                      _card("CUNIT1", "deg"), _card("CUNIT2", "deg"), _card("FREQ", float(freq)), _card("MJD", float(mjd)),
                      _card("BUNIT", "JY/PIXEL"), _card("STOKES", st), _card("TELESCOP", "KerrSplat")]
            data = (img[:, :, k] * omega / JY).T.astype(">f8")          # (y, x) in memory: x fastest on disk
            _write_hdu(f, cards, data)
    return path


def read_stokes_movie(paths, M_solar, D_pc, sigma, mjd0=None, freq=None):
    """read_stokes_movie(paths; M_solar, D_pc, mjd0, sigma, freq) (fits.jl:42-81): a StokesMovie from ehtim-style
    FITS images (one file per frame and frequency), grouped by MJD and FREQ, intensities from Jy/pixel to cgs, times
    to M from mjd0 (default the earliest frame), the pixel grid as a Camera in M; sigma per Stokes parameter in
    Jy/pixel. Returns (movie, camera, L)."""
    from .fit import StokesMovie
    L = gravitational_radius(M_solar)
    D = D_pc * PC
    frames = [(*read_stokes_fits(p), p) for p in paths]
    h1 = frames[0][1]
    psize_rad = h1["psize_deg"] * math.pi / 180
    nx, ny = frames[0][0].shape[:2]
    if h1["bunit"] != "JY/PIXEL":
        raise ValueError(f"unsupported BUNIT {h1['bunit']}")
    tojy = JY / psize_rad ** 2
    mjds = [fr[1]["mjd"] for fr in frames]
    freqs = []
    for fr in frames:
        if math.isnan(fr[1]["freq"]):
            if freq is None:
                raise ValueError(f"no FREQ in {fr[2]}")
            freqs.append(float(freq))
        else:
            freqs.append(fr[1]["freq"])
    t0 = min(mjds) if mjd0 is None else mjd0
    times, nus = sorted(set(mjds)), sorted(set(freqs))
    data = np.zeros((nx, ny, len(times), len(nus), 4))
    filled = np.zeros((len(times), len(nus)), dtype=bool)
    for fr, mjd, nu in zip(frames, mjds, freqs):
        if fr[0].shape[:2] != (nx, ny):
            raise ValueError("image sizes differ across the movie")
        k, l = times.index(mjd), nus.index(nu)
        data[:, :, k, l, :] = fr[0] * tojy
        filled[k, l] = True
    if not filled.all():
        raise ValueError("the movie does not cover every (time, frequency) pair")
    tM = (np.asarray(times) - t0) * 86400 / time_unit(M_solar)
    psize_M = psize_rad * D / L
    al = np.array([(i + 1 - (nx + 1) / 2) * psize_M for i in range(nx)])
    be = np.array([(j + 1 - (ny + 1) / 2) * psize_M for j in range(ny)])
    cam = Camera.points(np.repeat(al[:, None], ny, 1).ravel(order="F"), np.repeat(be[None, :], nx, 0).ravel(order="F"), (nx, ny))
    movie = StokesMovie(data, tM, nus, np.asarray(sigma, dtype=np.float64) * tojy)
    return movie, cam, L


# ---- uvfits (uvfits.jl:1-205) ----------------------------------------------------------------------------------------
class Observation:
    """Visibilities read from a uvfits file (uvfits.jl:9-36): per row time (UT hours of day mjd), tint (s), 0-based
    station indices s1, s2 into stations, u, v (wavelengths, as stored), the Stokes visibilities vis (complex
    (n, 4), Jy) and their noise sigma ((n, 4), Jy; inf where a row lacks the Stokes parameter), the circular products
    coh (RR, LL, RL, LR) and their noise; freq, bandwidth (Hz), ra (hours), dec (degrees), mjd and source."""

    def __init__(self, **kw):
        self.__dict__.update(kw)

    def __len__(self):
        return len(self.u)


def _bintable(h, data):
    """The columns of a binary table (name -> array); string fields as python strings."""
    nrow, rowlen = int(h["NAXIS2"]), int(h["NAXIS1"])
    cols, off = {}, 0
    types = {"L": ("?", 1), "X": ("u1", 1), "B": ("u1", 1), "I": (">i2", 2), "J": (">i4", 4), "K": (">i8", 8), "A": ("S1", 1),
             "E": (">f4", 4), "D": (">f8", 8), "C": (">c8", 8), "M": (">c16", 16)}
    fields = []
    for i in range(1, int(h["TFIELDS"]) + 1):
        form = str(h[f"TFORM{i}"]).strip()
        digits = "".join(ch for ch in form if ch.isdigit())
        code = form[len(digits):][:1]
        rep = int(digits) if digits else 1
        dt, size = types[code]
        fields.append((str(h.get(f"TTYPE{i}", f"COL{i}")).strip().upper(), code, rep, dt, off))
        off += rep * size
    raw = np.frombuffer(data[:nrow * rowlen], dtype=np.uint8).reshape(nrow, rowlen)
    for name, code, rep, dt, o in fields:
        size = types[code][1]
        chunk = raw[:, o:o + rep * size].copy()
        if code == "A":
            cols[name] = [bytes(r).decode("ascii", "replace").strip().strip("\x00") for r in chunk]
        else:
            arr = chunk.view(dt).reshape(nrow, rep) if rep > 1 else chunk.view(dt).reshape(nrow)
            cols[name] = arr.astype(arr.dtype.newbyteorder("="))
    return cols


def _products_from_stokes(vis, sig):
    sP = np.sqrt(sig[0] ** 2 + sig[3] ** 2)
    sX = np.sqrt(sig[1] ** 2 + sig[2] ** 2)
    return np.array([vis[0] + vis[3], vis[0] - vis[3], vis[1] + 1j * vis[2], vis[1] - 1j * vis[2]]), np.array([sP, sP, sX, sX])


def read_uvfits(path):
    """read_uvfits(path) (uvfits.jl:77-204): the random-groups primary HDU (AIPS layout, as ehtim writes): UU, VV scaled
    to wavelengths by the antenna table's reference frequency (else the FREQ axis), baselines decoded as 256 a1 + a2,
    the DATE parameters summed to a Julian date, the correlation products averaged over channels and IFs where their
    weights are positive and finite (noise sqrt(sum 1/w) / n), circular products converted to Stokes I = (RR + LL)/2,
    Q = (RL + LR)/2, U = (RL - LR)/(2i), V = (RR - LL)/2 with the noise in quadrature; rows without parallel hands
    dropped, one parallel hand taken as I without V, missing cross hands leaving Q, U with infinite noise."""
    hdus = read_hdus(path)
    hdr, raw = hdus[0]
    if hdr.get("GROUPS") is not True:
        raise ValueError(f"{path} is not a random-groups uvfits file")
    ngroups, npar, naxis = int(hdr["GCOUNT"]), int(hdr["PCOUNT"]), int(hdr["NAXIS"])
    dims = [int(hdr[f"NAXIS{i}"]) for i in range(2, naxis + 1)]
    ctypes = [str(hdr.get(f"CTYPE{i}", "")).strip().upper() for i in range(2, naxis + 1)]
    ptypes = [str(hdr[f"PTYPE{i}"]).strip().upper() for i in range(1, npar + 1)]
    pscal = np.array([float(hdr.get(f"PSCAL{i}", 1.0)) for i in range(1, npar + 1)])
    pzero = np.array([float(hdr.get(f"PZERO{i}", 0.0)) for i in range(1, npar + 1)])
    if not (ctypes[0] == "COMPLEX" and ctypes[1] == "STOKES" and "FREQ" in ctypes):
        raise ValueError(f"unsupported axis order {ctypes} in {path}")
    if dims[0] != 3:
        raise ValueError(f"expected (re, im, weight) triples on the COMPLEX axis of {path}")
    ifreq = ctypes.index("FREQ")
    iif = ctypes.index("IF") if "IF" in ctypes else None
    nstokes, nchan = dims[1], dims[ifreq]
    nif = dims[iif] if iif is not None else 1
    stokes0 = int(round(float(hdr[f"CRVAL{3}"])))
    dstokes = int(round(float(hdr.get(f"CDELT{3}", 1.0))))
    codes = [stokes0 + k * dstokes for k in range(nstokes)]      # -1..-4 = RR, LL, RL, LR; 1..4 = I, Q, U, V
    circular = all(c < 0 for c in codes)
    if not circular and not all(c > 0 for c in codes):
        raise ValueError(f"mixed STOKES axis {codes} in {path}")
    slot = lambda code: codes.index(code) if code in codes else None
    ra, dec = math.nan, math.nan
    for i, ct in enumerate(ctypes):
        if ct == "RA":
            ra = float(hdr[f"CRVAL{i + 2}"]) * 12 / 180
        if ct == "DEC":
            dec = float(hdr[f"CRVAL{i + 2}"])
    source = str(hdr.get("OBJECT", "")).strip()
    ch1 = float(hdr[f"CRVAL{ifreq + 2}"])
    bandwidth = abs(float(hdr.get(f"CDELT{ifreq + 2}", math.nan)))
    stations, nosta, rf = [], [], ch1
    for h, data in hdus[1:]:
        if str(h.get("EXTNAME", "")).strip().upper() == "AIPS AN":
            t = _bintable(h, data)
            stations = [s.strip() for s in t["ANNAME"]]
            nosta = [int(x) for x in t["NOSTA"]]
            if "FREQ" in h:
                rf = float(h["FREQ"])
    if not stations:
        raise ValueError(f"no AIPS AN table in {path}")
    order = np.argsort(nosta, kind="stable")
    stations = [stations[i] for i in order]
    nosta = [nosta[i] for i in order]
    pos = {n: i for i, n in enumerate(nosta)}
    iu = next(i for i, p in enumerate(ptypes) if p.startswith("UU"))
    iv = next(i for i, p in enumerate(ptypes) if p.startswith("VV"))
    ibl = ptypes.index("BASELINE")
    idates = [i for i, p in enumerate(ptypes) if p == "DATE"]
    itint = ptypes.index("INTTIM") if "INTTIM" in ptypes else None
    nblock = int(np.prod(dims))
    groups = np.frombuffer(raw, dtype=_dtype(hdr["BITPIX"]), count=ngroups * (npar + nblock)).astype(np.float64).reshape(ngroups, npar + nblock)
    pars = groups[:, :npar] * pscal + pzero
    block = groups[:, npar:]
    stride_chan = 3 * nstokes
    stride_if = 3 * nstokes * int(np.prod(dims[2:iif])) if iif is not None else 0
    rows = {k: [] for k in ("time", "tint", "s1", "s2", "u", "v", "vis", "sigma", "coh", "sigma_coh", "jd")}
    for g in range(ngroups):
        mean = np.full(nstokes, np.nan + 0j)
        sig = np.full(nstokes, np.inf)
        for p in range(nstokes):
            acc, var, n = 0j, 0.0, 0
            for m in range(nif):
                for c in range(nchan):
                    base = 3 * p + stride_chan * c + stride_if * m
                    w = block[g, base + 2]
                    if not (np.isfinite(w) and w > 0):
                        continue
                    acc += complex(block[g, base], block[g, base + 1])
                    var += 1.0 / w
                    n += 1
            if n:
                mean[p], sig[p] = acc / n, math.sqrt(var) / n
        if circular:
            rr, ll, rl, lr = slot(-1), slot(-2), slot(-3), slot(-4)
            hasrr = rr is not None and np.isfinite(sig[rr])
            hasll = ll is not None and np.isfinite(sig[ll])
            if not (hasrr or hasll):
                continue
            if hasrr and hasll:
                I, V = (mean[rr] + mean[ll]) / 2, (mean[rr] - mean[ll]) / 2
                sI = math.sqrt(sig[rr] ** 2 + sig[ll] ** 2) / 2
                sV = sI
            else:
                I, V = (mean[rr], 0j) if hasrr else (mean[ll], 0j)
                sI, sV = (sig[rr] if hasrr else sig[ll]), math.inf
            cross = rl is not None and lr is not None and np.isfinite(sig[rl]) and np.isfinite(sig[lr])
            Q = (mean[rl] + mean[lr]) / 2 if cross else 0j
            U = (mean[rl] - mean[lr]) / 2j if cross else 0j
            sQ = math.sqrt(sig[rl] ** 2 + sig[lr] ** 2) / 2 if cross else math.inf
            vis, sg = np.array([I, Q, U, V]), np.array([sI, sQ, sQ, sV])
            prod = lambda j: (0j, math.inf) if j is None or not np.isfinite(sig[j]) else (mean[j], sig[j])
            pr = [prod(j) for j in (rr, ll, rl, lr)]
            coh, scoh = np.array([x[0] for x in pr]), np.array([x[1] for x in pr])
        else:
            si = slot(1)
            if si is None or not np.isfinite(sig[si]):
                continue
            comp = lambda k: (0j, math.inf) if slot(k) is None or not np.isfinite(sig[slot(k)]) else (mean[slot(k)], sig[slot(k)])
            (Q, sQ), (U, sU), (V, sV) = comp(2), comp(3), comp(4)
            vis, sg = np.array([mean[si], Q, U, V]), np.array([sig[si], sQ, sU, sV])
            coh, scoh = _products_from_stokes(vis, sg)
        bl = int(math.floor(pars[g, ibl]))
        a1, a2 = bl // 256, bl - 256 * (bl // 256)
        if a1 not in pos or a2 not in pos:
            raise ValueError(f"baseline refers to a station absent from the AN table ({a1}, {a2})")
        rows["s1"].append(pos[a1]); rows["s2"].append(pos[a2])
        rows["u"].append(pars[g, iu] * rf); rows["v"].append(pars[g, iv] * rf)
        rows["jd"].append(float(sum(pars[g, i] for i in idates)))
        rows["tint"].append(pars[g, itint] if itint is not None else 0.0)
        rows["vis"].append(vis); rows["sigma"].append(sg); rows["coh"].append(coh); rows["sigma_coh"].append(scoh)
    if not rows["jd"]:
        raise ValueError(f"no unflagged parallel-hand data in {path}")
    jd = np.array(rows["jd"])
    mjd = int(math.floor(jd.min() - 2400000.5))
    return Observation(time=(jd - 2400000.5 - mjd) * 24, tint=np.array(rows["tint"]), s1=np.array(rows["s1"]), s2=np.array(rows["s2"]),
                       stations=stations, u=np.array(rows["u"]), v=np.array(rows["v"]), vis=np.array(rows["vis"]),
                       sigma=np.array(rows["sigma"]), coh=np.array(rows["coh"]), sigma_coh=np.array(rows["sigma_coh"]),
                       freq=rf, bandwidth=bandwidth, ra=ra, dec=dec, mjd=mjd, source=source)
