/*
 * syngen.h -- the drop-in C-ABI of libsyngen.so.
 *
 * The same 51 entry points, types and ownership rules as the reference's
 * src/core/extern.h:1-99, which its Python front end binds with ctypes
 * (lib/syngen.py:90-175).  Behind them the simulation runs on one MI355X GPU
 * through the kernel ABI in sg_hip.h.
 *
 * Error behaviour (extern.cpp:145-279): entry points that can fail catch
 * every error and return NULL (handles, run), false (load/save) or a null
 * ARRAY ({0, VOID_POINTER, NULL, false}).  Unlike the reference,
 * create_network / create_environment / build_state also return NULL on an
 * invalid description instead of terminating the process, and
 * get_layer_data / get_connection_data return a null ARRAY instead of
 * dereferencing a null pointer.
 *
 * Ownership (extern.cpp:12-66, state.cpp:706-764): get_neuron_data and
 * get_weight_matrix return writable, non-owning views of the host copy of
 * the state in the reference's host layout, valid until the next run(),
 * load_state() or destroy_state(); key listings are owned arrays released
 * with free_array / free_array_deep; run() returns a new report PROPS the
 * caller releases with destroy_properties().
 */
#ifndef SYNGEN_H
#define SYNGEN_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* extern.h:6-12 */
typedef enum {
    FLOAT_POINTER = 1,
    INT_POINTER = 2,
    STRING_POINTER = 3,
    PROPS_POINTER = 4,
    VOID_POINTER = 5
} POINTER_TYPE;

/* extern.h:14-19 (returned by value) */
typedef struct ARRAY {
    int size;
    POINTER_TYPE type;
    void *data;
    bool owner;
} ARRAY;

/* extern.h:27-30 */
typedef void *NETWORK;
typedef void *ENVIRONMENT;
typedef void *STATE;
typedef void *PROPS;

/* extern.h:24-25 */
void free_array(ARRAY arr);
void free_array_deep(ARRAY arr);

/* Property trees -- extern.h:32-47 */
PROPS create_properties(void);
void add_property(PROPS properties, char *key, char *val);
void add_child(PROPS properties, char *key, PROPS child);
void add_to_array(PROPS properties, char *key, char *val);
void add_to_child_array(PROPS properties, char *key, PROPS props);
ARRAY get_keys(PROPS properties);
ARRAY get_child_keys(PROPS properties);
ARRAY get_array_keys(PROPS properties);
ARRAY get_child_array_keys(PROPS properties);
const char *get_property(PROPS properties, char *key);
PROPS get_child(PROPS properties, char *key);
ARRAY get_array(PROPS properties, char *key);
ARRAY get_child_array(PROPS properties, char *key);

/* Environments and networks -- extern.h:49-55 */
ENVIRONMENT create_environment(PROPS properties);
ENVIRONMENT load_env(char *filename);
bool save_env(ENVIRONMENT env, char *filename);
NETWORK create_network(PROPS properties);
NETWORK load_net(char *filename);
bool save_net(NETWORK network, char *filename);

/* State -- extern.h:57-67 */
STATE build_state(NETWORK net);
STATE build_load_state(NETWORK net, char *filename);
bool load_state(STATE state, char *filename);
bool save_state(STATE state, char *filename);
ARRAY get_neuron_data(STATE state, char *structure_name, char *layer_name,
                      char *key);
ARRAY get_layer_data(STATE state, char *structure_name, char *layer_name,
                     char *key);
ARRAY get_connection_data(STATE state, char *conn_name, char *key);
ARRAY get_weight_matrix(STATE state, char *conn_name, char *key);

/* Engine -- extern.h:69 */
PROPS run(NETWORK net, ENVIRONMENT env, STATE state, PROPS args);

/* Destruction -- extern.h:71-75 */
void destroy(void *obj);
void destroy_network(NETWORK net);
void destroy_environment(ENVIRONMENT env);
void destroy_state(STATE state);
void destroy_properties(PROPS props);

/* Devices -- extern.h:77-80.  GPUs are 0..n-1 and the host is n.
 * get_num_gpus is declared but not defined by the reference; it is defined
 * here. */
int get_num_gpus(void);
int get_cpu(void);
ARRAY get_gpus(void);
ARRAY get_all_devices(void);

/* Logging -- extern.h:82-84 */
void set_suppress_output(bool val);
void set_warnings(bool val);
void set_debug(bool val);

/* extern.h:86 */
void interrupt_engine(void);

/* Callbacks by name -- extern.h:88-92; addr is a function address:
 * io (int id, int size, void *data), weight (int id, int size, float *w),
 * indices (id, size, w, from_indices, to_indices, used),
 * distance (id, size, w, distances), delay (id, size, w, delays). */
void add_io_callback(char *name, long long addr);
void add_weight_callback(char *name, long long addr);
void add_indices_weight_callback(char *name, long long addr);
void add_distance_weight_callback(char *name, long long addr);
void add_delay_weight_callback(char *name, long long addr);

/* MPI -- extern.h:94-97: one process, rank 0 of 1 (the reference's
 * behaviour without -D__MPI__, mpi_wrap.h:25-46). */
void initialize_mpi(void);
int get_mpi_rank(void);
int get_mpi_size(void);
void finalize_mpi(void);

#ifdef __cplusplus
}
#endif

#endif
