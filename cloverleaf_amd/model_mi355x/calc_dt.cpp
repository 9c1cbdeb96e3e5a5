// void calc_dt(global_variables&, int tile, double& local_dt, std::string& local_control,
//              double& xl_pos, double& yl_pos, int& jldt, int& kldt)   driver/calc_dt.h:26-27
#include "calc_dt.h"

#include <iostream>

#include "mi355x_model.h"

void calc_dt(global_variables &globals, int tile, double &local_dt, std::string &local_control, double &xl_pos, double &yl_pos, int &jldt,
             int &kldt) {
  (void)xl_pos;
  (void)yl_pos;  // never assigned by the reference either (src/serial/calc_dt.cpp:116-135)
  int l_control = 0, small = 0;
  const global_config &cfg = globals.config;
  // the recorded ideal_gas + viscosity + halos run with calc_dt as one launch
  if (!mi355x::deferred_calc_dt(globals, tile, local_dt, l_control, jldt, kldt, small))
    mi355x::check(clv_calc_dt(mi355x::chunk(globals, tile), cfg.dtmin, cfg.dtc_safe, cfg.dtu_safe, cfg.dtv_safe, cfg.dtdiv_safe, &local_dt,
                              &l_control, &jldt, &kldt, &small),
                  "calc_dt");
  if (l_control == 1) local_control = "sound";
  if (l_control == 2) local_control = "xvel";
  if (l_control == 3) local_control = "yvel";
  if (l_control == 4) local_control = "div";
  if (small != 0) std::cout << "Timestep information:" << std::endl << "j, k                 : " << jldt << " " << kldt << std::endl
                            << "timestep : " << local_dt << std::endl;
}
