"""Shallow direct-render golden cases (render_fractal_ex[_planes] with
use_perturbation = 0): all four formulas, both coordinate and bailout
conventions, plane hints and an output bias."""

BASE_X = "-0.743643887037158704752191506114774"
BASE_Y = "0.131825904205311970493132056385139"

CASES = [
    dict(name="direct_mandel_full", width=192, height=108, zoom="1", x="-0.5", y="0", max_iter=400,
         formula=0, planes=True, options={}),
    dict(name="direct_mandel_kalles", width=160, height=90, zoom="2000", x=BASE_X, y=BASE_Y, max_iter=3000,
         formula=0, planes=True, options=dict(escape_radius_mode=1, coordinate_mode=1)),
    dict(name="direct_mandel_seahorse", width=160, height=90, zoom="50000", x="-0.7453", y="0.1127",
         max_iter=5000, formula=0, planes=False, options=dict(output_bias=4000.0)),
    dict(name="direct_mandel_hints", width=128, height=72, zoom="30", x="-0.1011", y="0.9563", max_iter=2000,
         formula=0, planes=True, options=dict(plane_flags=3)),
    dict(name="direct_julia", width=160, height=90, zoom="1.2", x="0", y="0", max_iter=1000, formula=1,
         julia_re=-0.8, julia_im=0.156, planes=True, options={}),
    dict(name="direct_julia_kalles", width=128, height=72, zoom="7", x="0.1", y="-0.2", max_iter=1500,
         formula=1, julia_re=-0.7269, julia_im=0.1889, planes=False,
         options=dict(escape_radius_mode=1, coordinate_mode=1)),
    dict(name="direct_burning_ship", width=160, height=90, zoom="25", x="-1.762", y="-0.028", max_iter=800,
         formula=2, planes=False, options={}),
    dict(name="direct_tricorn", width=160, height=90, zoom="1.5", x="-0.3", y="0", max_iter=600, formula=3,
         planes=False, options=dict(escape_radius_mode=1)),
]
