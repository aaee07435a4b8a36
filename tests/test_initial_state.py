"""CPU pins of the reference's fresh start (initfd with gfortran's random
numbers, setup/state/index.ts:12-86, shared/random/gfortran/index.ts:8-51)
and of the 365-day (Gregorian) calendar (calendar/date/index.ts:10-121),
through the host set-up of a host-only engine (no GPU).
"""
import calendar
import datetime

import numpy as np
import pytest

# libgfortran's random_seed scrambling keys as the reference holds them
# (shared/random/gfortran/index.ts:10-15)
KEYS = [0xBD0C5B6E50C2DF49, 0xD46061CD46E1DF38, 0xBB4F4D4ED6103544, 0x114A583D0756AD39]
# xoshiro256** from the state {1, 2, 3, 4}: the published known-answer
# sequence of Blackman and Vigna's reference implementation
XOSHIRO_1234 = [11520, 0, 1509978240, 1215971899390074240, 1216172134540287360,
                607988272756665600, 16172922978634559625, 8476171486693032832,
                10595114339597558777, 2904607092377533576]
SCALE = 1e-4 / np.sqrt(2.0)


def seed_for_state(state):
    """The eight SEED words whose random_seed(put=) state is `state`: each
    state word is two 32-bit seed words, the seed read in reverse order,
    XORed with the key."""
    rev = []
    for i, s in enumerate(state):
        w = s ^ KEYS[i]
        rev += [w & 0xFFFFFFFF, w >> 32]
    words = rev[::-1]
    return [w - (1 << 32) if w >= 1 << 31 else w for w in words]


def fresh(exo, **kw):
    cfg = exo.default_config(initial_state=1, **{k: v for k, v in kw.items() if k != "seed"})
    if "seed" in kw:
        for i, w in enumerate(kw["seed"]):
            cfg.seed[i] = w
    return cfg


def initial_fields(e, m, names):
    ini = e.read_table(m, "initial")
    lay = {n: (o, c) for n, o, c in e.layout(m)}
    return {n: ini[lay[n][0]:lay[n][0] + lay[n][1]] for n in names}


def test_kick1_noise_is_gfortran_random_number(exo):
    """KICK 1 adds 1e-4/sqrt(2) (u - 1/2) to every non-zonal coefficient of
    ln(ps) in storage order, u the top 53 bits of xoshiro256** times 2^-53:
    with the seed whose scrambled state is {1, 2, 3, 4} the first draws are
    the published sequence."""
    with exo.Engine(-1, 1) as e:
        m = e.add_member(fresh(exo, seed=seed_for_state([1, 2, 3, 4])))
        T = e.dims(m)[0]
        f = initial_fields(e, m, ["sp", "spm"])
        sp = f["sp"]
        ntp1 = T + 1
        assert np.all(sp[:2 * ntp1] == 0)  # zonal modes untouched
        want = np.array([SCALE * ((x >> 11) * 2.0 ** -53 - 0.5) for x in XOSHIRO_1234])
        assert np.array_equal(sp[2 * ntp1:2 * ntp1 + 10], want)
        assert np.array_equal(f["spm"], sp)
        rest = sp[2 * ntp1:]
        assert np.all(np.abs(rest) <= SCALE / 2) and np.all(rest != 0)
        # white noise: mean near zero, variance of a uniform draw
        assert abs(rest.mean()) < 4 * SCALE / np.sqrt(12 * rest.size)
        assert rest.std() == pytest.approx(SCALE / np.sqrt(12), rel=0.15)


def test_seed_from_the_clock(exo):
    """initrandom: a SEED whose first word is zero takes (clock + 37 i)
    truncated to 32 bits, i = 1..8."""
    clock = 1796400000123
    words = [(clock + 37 * (i + 1)) % (1 << 32) for i in range(8)]
    words = [w - (1 << 32) if w >= 1 << 31 else w for w in words]
    with exo.Engine(-1, 3) as e:
        a = e.add_member(fresh(exo, seed=[0] * 8, clock_ms=clock))
        b = e.add_member(fresh(exo, seed=words))
        c = e.add_member(fresh(exo, seed=[0] * 8, clock_ms=clock + 1))
        sa, sb, sc = (initial_fields(e, x, ["sp"])["sp"] for x in (a, b, c))
        assert np.array_equal(sa, sb)
        assert not np.array_equal(sa, sc)


def test_kick_variants(exo):
    """KICK 2: one draw for both parts of each mode with n + m even and m > 0;
    KICK 3: mode (1, 1) alone; KICK 0: no noise."""
    with exo.Engine(-1, 3) as e:
        m2 = e.add_member(fresh(exo, kick=2))
        m3 = e.add_member(fresh(exo, kick=3))
        m0 = e.add_member(fresh(exo, kick=0))
        T = e.dims(m2)[0]
        sp2 = initial_fields(e, m2, ["sp"])["sp"].reshape(-1, 2)
        r = 0
        for m in range(T + 1):
            for n in range(m, T + 1):
                if m > 0 and (n + m) % 2 == 0:
                    assert sp2[r, 0] == sp2[r, 1] != 0
                    assert abs(sp2[r, 0]) <= SCALE / 2
                else:
                    assert sp2[r, 0] == sp2[r, 1] == 0
                r += 1
        sp3 = initial_fields(e, m3, ["sp"])["sp"]
        w11 = 2 * (T + 1) + 2  # mode (m=1, n=1) follows the T+1 zonal modes
        want = np.zeros_like(sp3)
        want[w11], want[w11 + 1] = SCALE, SCALE / 2
        assert np.array_equal(sp3, want)
        assert np.all(initial_fields(e, m0, ["sp"])["sp"] == 0)


@pytest.mark.parametrize("trunc,levels", [(21, 10), (42, 5)])
def test_initfd_fields(exo, trunc, levels):
    """initfd: each level's temperature is the global mean (mode 0) of
    setzt's restoration profile at that level, vorticity the solid-body
    mode plavor = 1/sqrt(3/8) at (m, n) = (0, 1), divergence and humidity
    zero, both time levels equal."""
    with exo.Engine(-1, 1) as e:
        m = e.add_member(fresh(exo, truncation=trunc, levels=levels))
        T, _, _, L = e.dims(m)
        nrsp = (T + 1) * (T + 2)
        f = initial_fields(e, m, ["st", "stm", "sz", "szm", "sd", "sdm", "sq", "sqm"])
        sr = e.read_table(m, "srp").reshape(L, nrsp)
        st = f["st"].reshape(L, nrsp)
        assert np.array_equal(st[:, 0], sr[:, 0])
        assert np.all(st[:, 1:] == 0)
        # the restoration profile's global mean is the profile's mean
        # temperature in model units: positive deviations are not required,
        # but it is finite and varies with height
        assert np.all(np.isfinite(sr[:, 0])) and np.ptp(sr[:, 0]) > 0
        sz = f["sz"].reshape(L, nrsp)
        assert np.all(sz[:, 2] == pytest.approx(1.63299310207, abs=0))
        sz[:, 2] = 0
        assert np.all(sz == 0)
        for name in ("sd", "sdm", "sq", "sqm"):
            assert np.all(f[name] == 0), name
        assert np.array_equal(f["stm"], f["st"])
        assert np.array_equal(f["szm"].reshape(L, nrsp)[:, 2], np.full(L, 1.63299310207))


def test_nspinit_takes_ln_ps_from_the_orography(exo):
    """NSPINIT > 0: ln(ps) = -so cv^2 / (R tgr) (setup/index.ts:443-446), the
    synthetic start's ln(ps) on the same terrain."""
    with exo.Engine(-1, 2) as e:
        a = e.add_member(fresh(exo, nspinit=1))
        b = e.add_member(exo.default_config())
        spa = initial_fields(e, a, ["sp"])["sp"]
        spb = initial_fields(e, b, ["sp"])["sp"]
        assert np.array_equal(spa, spb)
        so = e.read_table(a, "sop")
        k = e.constants(a, 0)
        assert spa[0] < 0 and so[0] > 0
        assert np.allclose(spa, -so * k["cv"] ** 2 / (k["gascon"] * 288.0), rtol=1e-12, atol=0)


def test_gregorian_dates(exo):
    """A 365-day year runs the proleptic Gregorian calendar: the date of a
    step is Python's date of the same day count (ordinal 1 = 1 January of
    year 1 = day 366 from the start of year 0); the run starts on 1 January
    of year 1."""
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(year_days=365))
        spd = 32  # 45-minute steps
        start = e.member_step(m)
        assert start == spd * 366
        assert e.member_date(m) == (1, 1, 1, 0, 0)
        rng = np.random.default_rng(7)
        steps = np.concatenate([start + rng.integers(0, spd * 365 * 9, 300),
                                [spd * (366 + 365 * 3 + 58), spd * (366 + 365 * 3 + 59),  # 28/29 Feb, year 4
                                 spd * (366 + 365 * 3 + 60), spd * 146097 * 5 + 17]])
        for s in steps.tolist():
            d = datetime.date.fromordinal(s // spd - 365)
            minutes = (s % spd) * 1440 // spd
            assert e.member_date_at(m, s) == (d.year, d.month, d.day, minutes // 60, minutes % 60), s


def test_gregorian_month_interpolation_is_continuous(exo):
    """The climatology weight between a month and its neighbour runs
    continuously through every month boundary and peaks at the month's
    middle, with each month's own length (calendar.monthrange): 29 days of
    February in a leap year, 28 otherwise."""
    with exo.Engine(-1, 1) as e:
        m = e.add_member(exo.default_config(year_days=365))
        spd = 32
        start = e.member_step(m)
        pos = []
        for s in range(start, start + spd * (365 * 3 + 366), 4):
            v = e.step_values(m, s)
            month, neighbour, weight = int(v[5]), int(v[6]), v[7]
            # the weight passes 1/2 on a month's last day: the boundary lies
            # at (days + 1) / (days + neighbour's days) of the way
            assert abs(neighbour - month) == 1 and 0 <= weight < 0.55
            pos.append(month + (weight if neighbour > month else -weight))
        pos = np.array(pos)
        # a coordinate in months that rises through each year and wraps at
        # the year's end: no jump within a year beyond one step's share
        d = np.diff(pos)
        d = d[np.abs(d) < 6]
        assert np.all(d >= -1e-12) and d.max() < 4 / (spd * 28) * 1.01
        # mid-February: the weight is zero at day (days + 1) / 2
        for year, y0 in ((1, 0), (4, 365 * 3)):
            days = calendar.monthrange(year, 2)[1]
            mid = spd * (366 + y0 + 31) + int(round(spd * ((days + 1) / 2 - 1)))
            v = e.step_values(m, mid)
            assert (int(v[5]), v[7]) == (2, pytest.approx(0.0, abs=1e-12)), year
