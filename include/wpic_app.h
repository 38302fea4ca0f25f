/*
 * wpic_app.h -- the reference's 2D problem drivers (2d/proj/weibel/app.f90
 * app__main :87-157, 2d/proj/reconnection/app.f90 app__main :87-150) on one
 * GPU: the reference's JSON config ("config" and "parameter" sections), its
 * derived parameters (init, weibel app.f90:266-364, reconnection
 * app.f90:260-357), the initial state or the restart file, the step loop and
 * its output cadence (io__ptcl, io__orb, moments + io__mom + energy.dat,
 * save_restart at max_elapsed and, in the Weibel app, after max_it) through wpic.h and
 * wpic_io.h.  The output files are the reference's (wpic_io.h).
 */
#ifndef WPIC_APP_H
#define WPIC_APP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the app described by the config file (NULL: "config.json", as the
 * reference's config_default) in WPIC_MODE_FAST or WPIC_MODE_PARITY. */
int wpic_app2d_main(const char *config, int mode);
/* The 3D Weibel app (3d/proj/weibel/app.f90 app__main :87-160, load_config,
 * init :266-380, energy_history :509-577 with one kinetic column per species)
 * the same way; "udrift" in the parameter section selects the
 * counter-streaming beams of the 3D Weibel workload. */
int wpic_app3d_main(const char *config, int mode);

#ifdef __cplusplus
}
#endif
#endif
