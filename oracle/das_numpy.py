"""Statement-by-statement numpy transliteration of the reference's MATLAB CPU path.

TEST INFRASTRUCTURE ONLY (tests/ use it to pin oracle.c; the product never
imports it).  Every block mirrors the vectorised MATLAB statements it cites,
including MATLAB's 1-based index arithmetic (converted at the point of use),
so that loop restructuring or an index slip in oracle.c shows up as a
mismatch.  Spectra come from numpy's FFT (pocketfft) rather than oracle.c's
own DFT, which pins that part independently as well.  Pure numpy: meant for
tiny grids only.
"""
from __future__ import annotations

import numpy as np


def deg2rad(d):
    return (np.pi / 180) * d


# +f_numbers/+distance/constant.m:90-121
def f_distance_constant(distance_deg, F_ub, r):
    r = np.asarray(r, dtype=float)
    sin_distance = np.sin(deg2rad(distance_deg))
    one_plus_cos = 1 + np.cos(deg2rad(distance_deg))
    with np.errstate(divide="ignore", invalid="ignore"):
        ind_lb = r >= 1 / one_plus_cos
        ind_ub = r < 1 / sin_distance
    ind_transition = ~ind_lb & (r >= 0.5)
    ind_possible = ind_lb & ind_ub
    values = np.zeros(r.shape)
    rt = r[ind_transition]
    values[ind_transition] = np.sqrt(1 / (1 / rt - 1) ** 2 - 1) / 2
    temp = r[ind_possible] ** 2
    values[ind_possible] = (np.sqrt(2 * one_plus_cos * temp - 1) + one_plus_cos * sin_distance * temp) / (
        2 - 2 * sin_distance ** 2 * temp)
    values[ind_possible] = np.minimum(values[ind_possible], F_ub)
    values[~ind_ub] = F_ub
    return values


# +f_numbers/+grating/grating.m:72-75 and angle_lb.m:44-75
def f_angle_lb(angle_lb_deg, F_number_ub, distance_deg, r):
    r = np.asarray(r, dtype=float)
    thresh = np.sin(deg2rad(angle_lb_deg))
    lower_bound = 1 / (1 + thresh)
    upper_bound = 1 / (1 / np.sqrt(1 + (2 * F_number_ub) ** 2) + thresh)
    ind_relevant = (r > lower_bound) & (r < upper_bound)
    ind_impossible = r >= upper_bound
    values = np.zeros(r.shape)
    values[ind_relevant] = np.sqrt(1 / (1 / r[ind_relevant] - thresh) ** 2 - 1) / 2
    values[ind_impossible] = F_number_ub
    if distance_deg >= 0:
        values = np.maximum(values, f_distance_constant(distance_deg, np.inf, r))
    return values


# +windows/*.m compute_samples_scalar of |t| (window.m:81)
def window_samples(name, param, t):
    t = np.abs(np.asarray(t, dtype=float))
    if name == "boxcar":
        return (t < 1).astype(float)
    if name == "hann":
        return (1 + np.cos(np.pi * t)) / 2
    if name == "triangular":
        return 1 - t
    if name == "tukey":
        frac_rect = 1 - param
        inside = t < 1
        s = inside.astype(float)
        taper = inside & (t > frac_rect)
        s[taper] = (1 + np.cos(np.pi * ((t[taper] - frac_rect) / param))) / 2
        return s
    raise ValueError(name)


def _normalize(samples, normalization):
    # +normalizations/on.m:60-61 (sum over the element dimension 2)
    if normalization == "off":
        return samples
    samples_norm = samples / samples.sum(axis=1, keepdims=True)
    if isinstance(normalization, tuple) and normalization[0] == "windowed":
        # +normalizations/windowed.m:81-86 (rows: valid frequencies, columns: elements)
        _, name, param = normalization
        lengths_over_two = np.sum(samples_norm > 0, axis=0) / 2
        with np.errstate(divide="ignore", invalid="ignore"):
            temp = window_samples(name, param, (np.arange(samples.shape[0])[:, None] - lengths_over_two)
                                  / lengths_over_two)
            samples_norm = samples_norm * temp
    return samples_norm


def das_pw(positions_x, positions_z, data_RF, f_s, steering_angle, element_width, element_pitch,
           c_0, f_bounds, index_t0, window, F_values_fn, normalization):
    """One steering angle.  data_RF: [N_t, N_el] (MATLAB layout).  Returns [N_z, N_x] complex.

    F_values_fn(element_pitch_over_lambda) -> F values (the F-number object)."""
    positions_x = np.asarray(positions_x, dtype=float)
    positions_z = np.asarray(positions_z, dtype=float)
    data_RF = np.asarray(data_RF, dtype=float)
    # das_pw.m:311-320
    N_elements = data_RF.shape[1]
    M_elements = (N_elements - 1) / 2
    element_width_over_two = element_width / 2
    positions_ctr_x = np.arange(-M_elements, M_elements + 0.5, 1.0) * element_pitch
    positions_lbs_x = positions_ctr_x - element_width_over_two
    positions_ubs_x = positions_ctr_x + element_width_over_two
    # das_pw.m:323-327
    e_steering_x = -np.sin(steering_angle)
    e_steering_z = np.cos(steering_angle)
    position_ctr_x_ref = np.sign(e_steering_x) * positions_ctr_x[0]
    # das_pw.m:330-340  [N_pos_x, N_elements]
    dist_lateral = positions_ctr_x[None, :] - positions_x[:, None]
    indicator_left = dist_lateral < 0
    t_in_plus_shift = (e_steering_x * (positions_x[None, :] - position_ctr_x_ref)
                       + e_steering_z * positions_z[:, None]) / c_0 + index_t0 / f_s
    t_sc_lateral_squared = (dist_lateral / c_0) ** 2
    t_sc_axial_squared = (positions_z / c_0) ** 2
    # das_pw.m:343
    N_samples_t_pad = int(np.ceil((np.sqrt((positions_ctr_x[-1] - positions_ctr_x[0]) ** 2 + positions_z[0] ** 2)
                                   - positions_z[0]) * f_s / c_0))
    # das_pw.m:350-362 (1-based indices_f)
    N_samples_t = data_RF.shape[0]
    N_points_dft = N_samples_t + N_samples_t_pad + 1 - ((N_samples_t + N_samples_t_pad) % 2)
    index_f_lb = int(np.ceil(f_bounds[0] * N_points_dft / f_s)) + 1
    index_f_ub = int(np.floor(f_bounds[1] * N_points_dft / f_s)) + 1
    indices_f = np.arange(index_f_lb, index_f_ub + 1)
    axis_f_bp = f_s * (indices_f - 1) / N_points_dft
    axis_omega_bp = 2 * np.pi * axis_f_bp
    # das_pw.m:368-377  [N_f, N_z]
    element_pitch_over_lambda = element_pitch * axis_f_bp / c_0
    F_number_values = F_values_fn(element_pitch_over_lambda)
    F_ub = np.sqrt(axis_f_bp[:, None] * positions_z[None, :] / (2.88 * c_0))
    F_number_values = np.minimum(F_number_values[:, None], F_ub)
    # das_pw.m:383-384
    data_RF_dft = np.fft.fft(data_RF, N_points_dft, axis=0)
    data_RF_dft_analy = 2 * data_RF_dft[indices_f - 1, :]
    # das_pw.m:387-390
    with np.errstate(divide="ignore"):
        width_aperture_over_two_desired = positions_z[None, :] / (2 * F_number_values)
    image = np.zeros((positions_z.size, positions_x.size), dtype=complex)
    for index_pos_x in range(positions_x.size):
        px = positions_x[index_pos_x]
        # das_pw.m:399-400 (1-based)
        with np.errstate(invalid="ignore"):
            lb = np.maximum(np.ceil(M_elements + (px - width_aperture_over_two_desired) / element_pitch) + 1, 1)
            ub = np.minimum(np.floor(M_elements + (px + width_aperture_over_two_desired) / element_pitch) + 1,
                            N_elements)
        indicator_valid = lb <= ub
        # das_pw.m:406-411
        wl = np.zeros(indicator_valid.shape)
        wl[indicator_valid] = px - positions_lbs_x[lb[indicator_valid].astype(int) - 1]
        wr = np.zeros(indicator_valid.shape)
        wr[indicator_valid] = positions_ubs_x[ub[indicator_valid].astype(int) - 1] - px
        indices_pos_z = np.nonzero(np.any(indicator_valid, axis=0))[0]
        for index_pos_z in indices_pos_z:
            valid = indicator_valid[:, index_pos_z]
            # das_pw.m:427-430
            t_sc = np.sqrt(t_sc_lateral_squared[index_pos_x, :] + t_sc_axial_squared[index_pos_z])
            tof = t_in_plus_shift[index_pos_z, index_pos_x] + t_sc
            weights = np.exp(1j * axis_omega_bp[valid][:, None] * tof[None, :])
            # das_pw.m:436-454
            if window[0] == "boxcar":
                with np.errstate(divide="ignore", invalid="ignore"):
                    ws = window_samples(window[0], window[1], dist_lateral[index_pos_x, :][None, :]
                                        / width_aperture_over_two_desired[valid, index_pos_z][:, None])
            else:
                left = indicator_left[index_pos_x, :]
                wsl = window_samples(window[0], window[1], dist_lateral[index_pos_x, left][None, :]
                                     / wl[valid, index_pos_z][:, None])
                wsr = window_samples(window[0], window[1], dist_lateral[index_pos_x, ~left][None, :]
                                     / wr[valid, index_pos_z][:, None])
                ws = np.concatenate([wsl, wsr], axis=1)
            # das_pw.m:457
            ws = _normalize(ws, normalization)
            # das_pw.m:462-467
            focused = np.sum(ws * weights * data_RF_dft_analy[valid, :], axis=1)
            image[index_pos_z, index_pos_x] = np.sum(focused, axis=0)
    return image
