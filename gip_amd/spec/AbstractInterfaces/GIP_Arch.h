// Node abstraction (interface of GIP_Arch.h:4-17 in the reference).
#ifndef _gip_arch_
#define _gip_arch_

class GIP_Arch {
   public:
    GIP_Arch() {}
    GIP_Arch(int cpus, int cores, int devs) : num_cpus(cpus), num_cores(cores), num_devs(devs) {}
    virtual ~GIP_Arch() {}
    virtual void postConstruct(int, int, int) {}
    virtual void setUp() {}

   protected:
    int num_cpus = 1;
    int num_cores = 1;
    int num_devs = 1;
};

#endif
