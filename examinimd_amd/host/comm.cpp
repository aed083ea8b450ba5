// comm.cpp -- Comm base (src/comm.cpp:40-69) and CommSerialHIP
// (src/comm_types/comm_serial.cpp:41-129).
#include "comm.h"

Comm::Comm(System* s, T_X_FLOAT comm_depth_) : system(s), comm_depth(comm_depth_) {}
Comm::~Comm() {}
void Comm::init() {}
void Comm::exchange() {}
void Comm::exchange_halo() {}
void Comm::update_halo() {}
const emd_halo* Comm::device_halo() const { return nullptr; }
void Comm::update_force() {}
void Comm::reduce_float(T_FLOAT*, T_INT) {}
void Comm::reduce_int(T_INT*, T_INT) {}
void Comm::reduce_max_float(T_FLOAT*, T_INT) {}
void Comm::reduce_max_int(T_INT*, T_INT) {}
void Comm::reduce_min_float(T_FLOAT*, T_INT) {}
void Comm::reduce_min_int(T_INT*, T_INT) {}
void Comm::scan_int(T_INT*, T_INT) {}
void Comm::weighted_reduce_float(T_FLOAT*, T_INT*, T_INT) {}
void Comm::create_domain_decomposition() {
  system->sub_domain_lo_x = 0.0;
  system->sub_domain_lo_y = 0.0;
  system->sub_domain_lo_z = 0.0;
  system->sub_domain_x = system->sub_domain_hi_x = system->domain_x;
  system->sub_domain_y = system->sub_domain_hi_y = system->domain_y;
  system->sub_domain_z = system->sub_domain_hi_z = system->domain_z;
}
int Comm::process_rank() { return 0; }
int Comm::num_processes() { return 1; }
void Comm::error(const char* errormsg) {
  printf("%s\n", errormsg);
  exit(1);
}
const char* Comm::name() { return "InvalidComm"; }

CommSerialHIP::CommSerialHIP(System* s, T_X_FLOAT comm_depth_) : Comm(s, comm_depth_) {
  memset(&h, 0, sizeof(h));
}
CommSerialHIP::~CommSerialHIP() { if (system && system->ctx) emd_halo_free(system->ctx, &h); }

void CommSerialHIP::exchange() {
  EMD_CALL(system->ctx, emd_comm_exchange(system->ctx, system->view()));
}

void CommSerialHIP::exchange_halo() {
  system->N_ghost = 0;
  emd_system* v = system->view();
  EMD_CALL(system->ctx, emd_comm_exchange_halo(system->ctx, v, comm_depth, &h));
  system->pull(v);
}

void CommSerialHIP::update_halo() {
  EMD_CALL(system->ctx, emd_comm_update_halo(system->ctx, system->view(), &h));
}

void CommSerialHIP::update_force() {
  EMD_CALL(system->ctx, emd_comm_update_force(system->ctx, system->view(), &h, system->precision));
}

const char* CommSerialHIP::name() { return "CommSerial"; }
